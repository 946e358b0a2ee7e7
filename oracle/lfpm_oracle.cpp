// lfpm_oracle.cpp — TEST INFRASTRUCTURE ONLY.
//
// Deterministic CPU restatement (FP64, with an FP32 twin) of the micro-image PatchMatch path of
// fengyunxuan6/plenoptic-for-depth, used as the parity checker for the HIP product library and as
// the CPU baseline ("port").  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load it.  It shares no code with the product: its own Philox, its own brute-force lens
// layout, its own NCC, solver, filter, refocus and fit.
//
// What each function follows (paths relative to the reference repository root):
//   philox()                : Philox4x32-10, Salmon et al., "Parallel Random Numbers: As Easy as
//                             1, 2, 3", SC'11 (replaces curand, MIStereo/AdaptMIPMFrame.cu:501).
//   layout()                : crop rule Common/Common.cpp:34-38, abandon rule
//                             DepthSolver.cpp:1880-1915, view order Common/QuadTree.h:137-148,
//                             neighbour sort by baseline MVStereo/LFDepthInfo.cpp:214-241
//                             (brute force over all MI pairs, ties broken by angle).
//   ncc()                   : ComputeBilateralNCC_Frame, MIStereo/AdaptMIPMFrame.cu:73-179, with
//                             ComputeBilateralWeight MIStereo/CudaUtil.cu:145-150.
//   cost_vector()           : ComputeMultiViewCostVector_Frame, AdaptMIPMFrame.cu:181-225
//                             (blur term off: the synthetic workloads carry no blur maps).
//   topk()                  : ComputeSelectedViewsTopK_Frame, AdaptMIPMFrame.cu:227-282.
//   weighted_cost()         : ComputeWeightedCost_Frame, AdaptMIPMFrame.cu:284-317.
//   random_normal()         : GenerateRandomNormal_MIPM, MIStereo/AdaptMIPMUtil.cu:463-486
//                             (Marsaglia 1972), rejection test done exactly in integers.
//   perturbed_normal()      : GeneratePerturbedNormalLF, MIStereo/CudaUtil.cu:105-142.
//   cross_view_proposal()   : BuildCrossViewPlaneProposal, AdaptMIPMFrame.cu:350-398, and
//                             EncodePlaneHypothesisFromDispPlaneDevice :36-53.
//   refine()                : PlaneHypothesisRefinement_Frame, AdaptMIPMFrame.cu:400-471, with
//                             the candidate set fixed to the ACMP intent of
//                             MIStereo/AdaptMIPMUtil.cu:181-183 (depths[] x normals[]).
//   init_pixel()            : RandomInitializationForFrame, AdaptMIPMFrame.cu:474-518.
//   update_pixel()          : CheckerboardUpdateFrame, AdaptMIPMFrame.cu:520-947.
//   solve()                 : RunPatchMatchCUDAForFrame_Impl, AdaptMIPMFrame.cu:949-995.
//   filter_f32()            : MIDisparityFilterCPU::FilterTile, CheckModule/MIDisparityFilter.cpp:
//                             423-693 (+ helpers :25-174) with snapshot reads (every read sees
//                             the pre-filter state), configuration MIStereoMatch.cpp:759-778.
//   orc_filter_f32_inplace(): the same filter in FilterGlobal's order (MIDisparityFilter.cpp:
//                             973-985): MIs in map order, each invalidating its rejects before
//                             the next one reads them.
//   refocus()               : projection u = (X - c)/v + c_mi of LFRefocus::BackProject2MI,
//                             FocalStack/LFRefocus.cpp:2100-2125 and :244-281.
//   fit_v2r()/v2r_convert() : VirtualToRealDepthFunc::BehavioralModel / convertVirtualToRealDepth,
//                             FocalStack/VirtualToRealDepthFunc.cpp:3390-3451 (Householder QR
//                             least squares in FP64 instead of cv::solve(DECOMP_SVD)).
//   fixed_neighbors()       : SelectNeighbors::FixedPositionAlgorithm / SelectFixedPosition-
//                             MIByMode1, MINeighbor/SelectNeighbors.cpp:1035-1352, sort of
//                             MVStereo/LFDepthInfo.cpp:431-453, garbage rule DepthSolver.cpp:
//                             2933-2946 (the reference's if/else chain, branch for branch).
//   solve_acmm()            : AdaptMIPMFrameACMM (MIStereo/AdaptMIPMFrameACMM.cpp:155-665, .cu:
//                             373-991): level schedule, INTER_AREA level images (area average),
//                             warm start with the 3x3 seed gate, warm-seeded init, detail
//                             restorer, top-k metadata after every update, post-refine pass.
//   disparity_mosaic()      : DepthSolver::WriteDisMap_TileKey_new_AccuSequence,
//                             DepthSolver.cpp:6716-6897 (float map instead of the renderings).
//   vd_map()                : DepthSolver::Virtual_depth_map_TileKeySequence,
//                             DepthSolver.cpp:7832-8000 (median splat, per-channel colour
//                             medians :7919-7927, colour fill, 10 x medianBlur 5x5 with
//                             replicated borders per OpenCV's documentation, each channel).
//
// Semantics the reference leaves undefined and that this restatement fixes (the product does
// the same): planes whose normal is more than acos(cos_max_slant) from the optical axis score the
// maximum cost 2 (cos_max_slant = 0 is the reference behaviour); hypothesis w is the disparity at the owning pixel (kept as the reference does);
// refinement uses the ACMP depth x normal combination; disp record = (d, |d|*|b|/base, |b|,
// |b|/(d_real+1e-7)) of the lowest-cost view; bilinear reads clamp to the crop edge; the
// cross-view proposal denominator 1 + a*du + b*dv is kept exactly as the reference writes it.
//
// Parity pins: tests/test_oracle_pins.py (Philox KATs, closed-form NCC cases, hex-ring closed
// forms, plane identities, exact-consistency filter scenes, constant-v refocus, exact-model fit).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

static const int KMAX = 32;

// Filter reject reasons (CheckModule/MIDisparityFilter.h:102-109).
enum : uint8_t { PR_KEEP = 0, PR_COST = 80, PR_NOSUPPORT = 140, PR_CYCLE = 200, PR_SPIKE = 255 };

// ------------------------------------------------------------------------------------------
// Philox4x32-10
// ------------------------------------------------------------------------------------------
static inline void philox(const uint32_t cin[4], const uint32_t kin[2], uint32_t out[4]) {
    uint32_t c0 = cin[0], c1 = cin[1], c2 = cin[2], c3 = cin[3];
    uint32_t k0 = kin[0], k1 = kin[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Per-(pixel, step) draw stream.  Slot s -> Philox block s/4, word s%4 of
// ctr = (pixel_uid, step, s/4, 0), key = (seed, frame_uid).  A stream built over a word array
// returns fixed[s] instead (the pin entry points feed hand-chosen draws through it).
struct Stream {
    uint32_t key[2];
    uint32_t pix, step;
    uint32_t cached_block;
    uint32_t words[4];
    const uint32_t* fixed = nullptr;
    Stream(uint32_t seed, uint32_t uid, uint32_t pix_, uint32_t step_)
        : pix(pix_), step(step_), cached_block(0xFFFFFFFFu) {
        key[0] = seed; key[1] = uid;
    }
    explicit Stream(const uint32_t* w) : pix(0), step(0), cached_block(0xFFFFFFFFu), fixed(w) {
        key[0] = key[1] = 0;
    }
    uint32_t raw(uint32_t slot) {
        if (fixed) return fixed[slot];
        const uint32_t b = slot >> 2;
        if (b != cached_block) {
            const uint32_t c[4] = {pix, step, b, 0u};
            philox(c, key, words);
            cached_block = b;
        }
        return words[slot & 3u];
    }
};

enum : uint32_t {
    SLOT_INIT_DEPTH = 0,
    SLOT_INIT_MARSAGLIA = 4,     // attempt a uses SLOT+2a, SLOT+2a+1
    SLOT_VIEW_DRAW = 0,          // draws 0..14
    SLOT_REF_DEPTH_RAND = 16,
    SLOT_REF_DEPTH_PERT = 17,
    SLOT_REF_ANGLE = 20,         // 20, 21, 22
    SLOT_REF_MARSAGLIA = 24,
    SLOT_WARM_JITTER = 40,       // ACMM warm-seed disparity jitter (init step)
    SLOT_WARM_ANGLE = 44,        // ACMM warm-seed normal perturbation, 44..46 (init step)
    MARSAGLIA_ATTEMPTS = 16,
    VIEW_DRAWS = 15,
};

template <typename T> static inline T uniform(uint32_t x) {
    // ((x >> 8) + 0.5) * 2^-24: exact in FP32 and FP64, in (0, 1)
    return ((T)(x >> 8) + (T)0.5) * (T)5.9604644775390625e-08;
}

// ------------------------------------------------------------------------------------------
// Parameters (field-for-field identical to lfpm_pm_params / lfpm_filter_params /
// lfpm_layout_params of the product ABI)
// ------------------------------------------------------------------------------------------
struct PmParams {
    int crop; float base; int radius; int radius_increment;
    float sigma_spatial; float sigma_color; float min_var; int top_k;
    float depth_min, depth_max; int iterations; int max_neighbors;
    uint32_t seed; uint32_t frame_uid; int storage;
    float cos_max_slant; float blur_lambda;
};

struct FilterParams {
    int max_triplet; float min_valid_disp; int use_selected_views_only;
    int clear_selected_views_when_invalid; int enable_cost_filter; float max_cost;
    int enable_cycle_geo_filter; float max_geo_err_px; int enable_cycle_photo_filter;
    float max_photo_err_u; int min_good_neighbors; int enable_spike_filter;
    float spike_abs_diff; int spike_min_neighbors;
};

struct LayoutParams {
    float pitch; int crop; int raw_w, raw_h; int max_neighbors; int lens_types;
    int neighbors_per_type[3];
};

// ------------------------------------------------------------------------------------------
// Solver
// ------------------------------------------------------------------------------------------
template <typename T> struct P4 { T x, y, z, w; };

template <typename T> struct Frame {
    int V, W, K;
    std::vector<T> img;           // V*W*W intensities in [0,1]
    std::vector<T> blur;          // V*W*W blur values in [0,1] (blur term on), else empty
    const float* centers;         // 2V
    const int* nbr;               // V*K
    const int* cnt;               // V
    PmParams p;
    T inv_base;

    T I(int v, int x, int y) const { return img[(size_t)v * W * W + (size_t)y * W + x]; }
    T Bv(int v, int x, int y) const { return blur[(size_t)v * W * W + (size_t)y * W + x]; }

    // Bilinear with clamp-to-edge of plane `src` (img or blur); (x, y) in [0, W) x [0, W).
    T bilinear_of(const std::vector<T>& src, int v, T x, T y) const {
        const T fx0 = std::floor(x), fy0 = std::floor(y);
        const int x0 = (int)fx0, y0 = (int)fy0;
        const int x1 = std::min(x0 + 1, W - 1), y1 = std::min(y0 + 1, W - 1);
        const T fx = x - fx0, fy = y - fy0;
        const size_t o = (size_t)v * W * W;
        const T a = src[o + (size_t)y0 * W + x0], b = src[o + (size_t)y0 * W + x1];
        const T c = src[o + (size_t)y1 * W + x0], d = src[o + (size_t)y1 * W + x1];
        const T top = a + fx * (b - a);
        const T bot = c + fx * (d - c);
        return top + fy * (bot - top);
    }
    T bilinear(int v, T x, T y) const { return bilinear_of(img, v, x, y); }
    // Magnitudes of the bilinear interpolant's partial derivatives at (x, y) (as bilinear_of).
    void bilinear_grad(const std::vector<T>& img, int v, T x, T y, double& gx, double& gy) const {
        const T fx0 = std::floor(x), fy0 = std::floor(y);
        const int x0 = (int)fx0, y0 = (int)fy0;
        const int x1 = std::min(x0 + 1, W - 1), y1 = std::min(y0 + 1, W - 1);
        const double fx = (double)(x - fx0), fy = (double)(y - fy0);
        const size_t o = (size_t)v * W * W;
        const double a = img[o + (size_t)y0 * W + x0], b = img[o + (size_t)y0 * W + x1];
        const double c = img[o + (size_t)y1 * W + x0], d = img[o + (size_t)y1 * W + x1];
        gx = std::fabs((1 - fy) * (b - a) + fy * (d - c));
        gy = std::fabs((1 - fx) * (c - a) + fx * (d - b));
    }
    int count(int v) const { return std::min(cnt[v], K); }
};

// Bilateral weights of one reference window (depend on the reference pixel only, so they are
// computed once per pixel and reused for every hypothesis and view).
template <typename T> struct Win {
    T w[(2 * 10 + 1) * (2 * 10 + 1)];
    T rc;
};

template <typename T>
static void window_weights(const Frame<T>& f, int v, int px, int py, Win<T>& ww) {
    const int r = f.p.radius, n = 2 * r + 1;
    ww.rc = f.I(v, px, py);
    const T ss = (T)2 * (T)f.p.sigma_spatial * (T)f.p.sigma_spatial;
    const T sc = (T)2 * (T)f.p.sigma_color * (T)f.p.sigma_color;
    for (int j = -r; j <= r; ++j)
        for (int i = -r; i <= r; ++i) {
            const int rx = px + i, ry = py + j;
            T wv = 0;
            if (rx >= 0 && rx < f.W && ry >= 0 && ry < f.W) {
                const T spatial = std::sqrt((T)(i * i + j * j));
                const T color = std::fabs(f.I(v, rx, ry) - ww.rc);
                wv = std::exp(-spatial / ss - color / sc);
            }
            ww.w[(j + r) * n + (i + r)] = wv;
        }
}

// How far one evaluation is from the places where FP32 and FP64 may legitimately disagree:
//  geo   smallest distance (px) by which the anchor or a window sample's source point would
//        have to move to change its inside-the-crop test (AdaptMIPMFrame.cu:104-108,120-131);
//        a coordinate whose step component is exactly 0 is an exact integer and never moves;
//  gate  smallest distance of the two window variances from the variance gate (:170-174), in
//        units of their estimated FP32 error (inf where the evaluation ends before the gate);
//  ferr  first-order estimate of the FP32 error of the cost: the moments summed in FP32 with
//        the window-centre shift (relative error n_eff * 2^-24 of each weighted absolute sum,
//        n_eff = ceil(n / 8) + 3 terms per partial sum, the reduction of 8 lanes included),
//        the source coordinates' FP32 rounding times the interpolant's slope, and the 1-ulp
//        reciprocal and reciprocal square root; propagated through the variance and covariance
//        into 1 - cov / sqrt(var_x var_y).  0 where the cost is exactly 2 by an early return.
struct NccMargins { double geo, gate, ferr; };

// Distance of (qx, qy) from a change of the inside test 0 <= q < W: for an inside point the
// nearest edge; for an outside point the largest violated edge distance (all of them must be
// crossed to get inside).  An exact axis (step component 0: q is an integer) is not counted.
template <typename T> static double inside_margin(T qx, T qy, T W, bool exact_x, bool exact_y) {
    const double inf = INFINITY;
    const double dx0 = exact_x ? inf : (double)qx, dx1 = exact_x ? inf : (double)W - (double)qx;
    const double dy0 = exact_y ? inf : (double)qy, dy1 = exact_y ? inf : (double)W - (double)qy;
    const bool in = qx >= (T)0 && qx < W && qy >= (T)0 && qy < W;
    if (in) return std::min(std::min(dx0, dx1), std::min(dy0, dy1));
    double m = 0;
    if (!exact_x && qx < (T)0) m = std::max(m, -dx0);
    if (!exact_x && qx >= W) m = std::max(m, -dx1);
    if (!exact_y && qy < (T)0) m = std::max(m, -dy0);
    if (!exact_y && qy >= W) m = std::max(m, -dy1);
    return m;
}

template <typename T>
static T ncc(const Frame<T>& f, const Win<T>& ww, int v, int k, int px, int py, const P4<T>& h,
             NccMargins* mg = nullptr) {
    if (mg) { mg->geo = INFINITY; mg->gate = INFINITY; mg->ferr = 0; }
    const int n = f.nbr[(size_t)v * f.K + k];
    if (n < 0) return (T)2;
    const T sx = ((T)f.centers[2 * v] - (T)f.centers[2 * n]) * f.inv_base;
    const T sy = ((T)f.centers[2 * v + 1] - (T)f.centers[2 * n + 1]) * f.inv_base;
    const T W = (T)f.W;
    const T ax = (T)px + sx * h.w, ay = (T)py + sy * h.w;
    const bool ex = sx == (T)0, ey = sy == (T)0;
    if (mg) mg->geo = inside_margin(ax, ay, W, ex, ey);
    if (!(ax >= (T)0 && ax < W && ay >= (T)0 && ay < W)) return (T)2;
    // slant gate: planes steeper than acos(cos_max_slant) are not scored
    if (!(std::fabs(h.z) >= (T)f.p.cos_max_slant)) return (T)2;
    const T A = -h.x / h.z, B = -h.y / h.z;
    const int r = f.p.radius, inc = f.p.radius_increment, nn = 2 * r + 1;
    T s_w = 0, s_x = 0, s_xx = 0, s_y = 0, s_yy = 0, s_xy = 0;
    T b_ref = 0, b_src = 0;  // blur sums (ComputeBilateralNCC_Frame :140-144)
    // margin bookkeeping (FP64): weighted absolute sums and the coordinate-rounding terms
    double e_ax = 0, e_ay = 0, e_axy = 0, e_axx = 0, e_ayy = 0, e_dy = 0, e_ydy = 0, e_xdy = 0;
    double e_b = 0, e_bd = 0;  // blur sums: absolute values about the centre blur, rounding
    int e_n = 0;
    const double u32 = 5.9604644775390625e-08;
    const bool blur = !f.blur.empty();
    for (int i = -r; i <= r; i += inc) {
        for (int j = -r; j <= r; j += inc) {
            const int rx = px + i, ry = py + j;
            if (rx < 0 || rx >= f.W || ry < 0 || ry >= f.W) continue;
            const T d = h.w + A * (T)i + B * (T)j;
            const T qx = (T)rx + sx * d, qy = (T)ry + sy * d;
            if (mg) mg->geo = std::min(mg->geo, inside_margin(qx, qy, W, ex, ey));
            if (!(qx >= (T)0 && qx < W && qy >= (T)0 && qy < W)) continue;
            if (blur) {
                b_ref += f.Bv(v, rx, ry);
                b_src += f.bilinear_of(f.blur, n, qx, qy);
            }
            const T wgt = ww.w[(j + r) * nn + (i + r)];
            const T x = f.I(v, rx, ry) - ww.rc;
            const T y = f.bilinear(n, qx, qy) - ww.rc;
            s_w += wgt;
            s_x += wgt * x;
            s_xx += wgt * x * x;
            s_y += wgt * y;
            s_yy += wgt * y * y;
            s_xy += wgt * x * y;
            if (mg) {
                // FP32 rounding of the source point: the step, slope and map products and sums
                // (a few roundings of terms up to |p| + |i| + |s| (|w| + |A i| + |B j|))
                const double dq = 8 * u32 * ((double)std::fabs((T)rx) + (double)std::fabs((T)ry) +
                                             (std::fabs((double)sx) + std::fabs((double)sy)) *
                                                 ((double)std::fabs(h.w) + std::fabs((double)A * i) +
                                                  std::fabs((double)B * j)));
                double gx, gy;
                f.bilinear_grad(f.img, n, qx, qy, gx, gy);
                if (blur) {
                    const double rbc = (double)f.Bv(v, px, py);
                    double bgx, bgy;
                    f.bilinear_grad(f.blur, n, qx, qy, bgx, bgy);
                    e_b += std::fabs((double)f.Bv(v, rx, ry) - rbc) +
                           std::fabs((double)f.bilinear_of(f.blur, n, qx, qy) - rbc);
                    e_bd += (ex ? 0 : bgx * dq) + (ey ? 0 : bgy * dq);
                }
                const double dy = (ex ? 0 : gx * dq) + (ey ? 0 : gy * dq) +
                                  4 * u32 * (std::fabs((double)y) + std::fabs((double)ww.rc));
                const double wd = (double)wgt, xd = (double)x, yd = (double)y;
                e_ax += wd * std::fabs(xd); e_ay += wd * std::fabs(yd);
                e_axx += wd * xd * xd; e_ayy += wd * yd * yd;
                e_axy += wd * std::fabs(xd * yd);
                e_dy += wd * dy; e_ydy += wd * std::fabs(yd) * dy; e_xdy += wd * std::fabs(xd) * dy;
                ++e_n;
            }
        }
    }
    // FP32 error estimates of the normalised moments (see NccMargins)
    double d_vx = 0, d_vy = 0, d_cov = 0;
    if (mg && s_w > (T)0) {
        const double sw = (double)s_w, g = (double)((e_n + 7) / 8 + 3) * u32;
        const double mxd = (double)s_x / sw, myd = (double)s_y / sw;
        const double ax = e_ax / sw, ay = e_ay / sw, axx = e_axx / sw, ayy = e_ayy / sw,
                     axy = e_axy / sw;
        // each normalised sum m = S / sw: |dm| <= g * (|S|-sum / sw + |m|) (sum and division)
        const double d_mx = g * (ax + std::fabs(mxd));
        const double d_my = g * (ay + std::fabs(myd)) + e_dy / sw;
        d_vx = g * (2 * axx) + 2 * std::fabs(mxd) * d_mx + u32 * (axx + mxd * mxd);
        d_vy = g * (2 * ayy) + 2 * e_ydy / sw + 2 * std::fabs(myd) * d_my + u32 * (ayy + myd * myd);
        d_cov = g * (2 * axy) + e_xdy / sw + std::fabs(mxd) * d_my + std::fabs(myd) * d_mx +
                u32 * (axy + std::fabs(mxd * myd));
    }
    if (s_w <= (T)1e-6) return (T)2;
    const T inv = (T)1 / s_w;
    const T mx = s_x * inv, my = s_y * inv;
    const T vx = s_xx * inv - mx * mx;
    const T vy = s_yy * inv - my * my;
    if (mg) {
        const double mv = (double)f.p.min_var;
        mg->gate = std::min(std::fabs((double)vx - mv) / std::max(d_vx, 1e-300),
                            std::fabs((double)vy - mv) / std::max(d_vy, 1e-300));
    }
    if (vx < (T)f.p.min_var || vy < (T)f.p.min_var) return (T)2;
    const T cov = s_xy * inv - mx * my;
    if (mg) {
        const double sq = std::sqrt((double)vx * (double)vy);
        const double rho = std::fabs((double)cov) / sq;
        mg->ferr = d_cov / sq + 0.5 * rho * (d_vx / (double)vx + d_vy / (double)vy) +
                   3 * u32 * rho;
    }
    T c = std::min((T)2, std::max((T)0, (T)1 - cov / std::sqrt(vx * vy)));
    if (blur) {
        // blur term of ComputeMultiViewCostVector_Frame (AdaptMIPMFrame.cu:211-218); the
        // early returns above are cost 2, which the clamp keeps at 2 with or without it
        const T db = b_ref - b_src;
        const T bw = std::exp((T)-0.007368f * db * db);
        c = std::min((T)2, std::max((T)0, c + (T)f.p.blur_lambda * ((T)1 - bw)));
        if (mg) {
            const double g = (double)((e_n + 7) / 8 + 3) * u32, kb = 0.007368;
            const double ddb = g * e_b + e_bd;
            mg->ferr += (double)f.p.blur_lambda * 2 * kb * std::fabs((double)db) *
                            std::exp(-kb * (double)db * (double)db) * ddb +
                        (double)f.p.blur_lambda * 4 * u32;
        }
    }
    return c;
}

template <typename T>
static void cost_vector(const Frame<T>& f, const Win<T>& ww, int v, int px, int py,
                        const P4<T>& h, T* cv) {
    for (int k = 0; k < KMAX; ++k) cv[k] = (T)2;
    const int c = f.count(v);
    for (int k = 0; k < c; ++k) cv[k] = ncc(f, ww, v, k, px, py, h);
}

template <typename T> static P4<T> disp_record(const Frame<T>& f, int v, int k, const P4<T>& h) {
    const int n = f.nbr[(size_t)v * f.K + k];
    const T bx = (T)f.centers[2 * v] - (T)f.centers[2 * n];
    const T by = (T)f.centers[2 * v + 1] - (T)f.centers[2 * n + 1];
    const T b = std::sqrt(bx * bx + by * by);
    const T d = h.w;
    const T dr = std::fabs(d) * b * f.inv_base;
    return P4<T>{d, dr, b, b / (dr + (T)1e-7)};
}

template <typename T>
static unsigned topk(const T* cv, int cnt, int top_k, T& mean, int& best) {
    best = -1;
    int valid = 0;
    for (int i = 0; i < cnt; ++i)
        if (cv[i] < (T)2) {
            ++valid;
            if (best < 0 || cv[i] < cv[best]) best = i;
        }
    if (valid <= 0) { mean = (T)2; return 0u; }
    const int k = std::min(valid, top_k);
    T tmp[KMAX];
    for (int i = 0; i < cnt; ++i) tmp[i] = cv[i];
    for (int i = 0; i < k; ++i)
        for (int j = i + 1; j < cnt; ++j)
            if (tmp[j] < tmp[i]) std::swap(tmp[i], tmp[j]);
    mean = 0;
    for (int i = 0; i < k; ++i) mean += tmp[i];
    mean /= (T)k;
    const T thr = tmp[k - 1];
    unsigned m = 0;
    for (int i = 0; i < cnt && i < 32; ++i)
        if (cv[i] <= thr) m |= 1u << i;
    return m;
}

template <typename T>
static T weighted_cost(const T* cv, const T* vw, int cnt, int& argmin) {
    T wn = 0, tot = 0, mc = (T)2;
    argmin = -1;
    for (int i = 0; i < cnt; ++i) {
        if (vw[i] > (T)0) { wn += vw[i]; tot += vw[i] * cv[i]; }
        if (cv[i] < mc) { mc = cv[i]; argmin = i; }
    }
    if (wn <= (T)1e-6) return (T)2;
    return tot / wn;
}

template <typename T> static void normalize3(T& x, T& y, T& z) {
    const T n = std::sqrt(x * x + y * y + z * z);
    x = x / n; y = y / n; z = z / n;
}

template <typename T> static void random_normal(Stream& st, uint32_t base, T n[3]) {
    int64_t A = 0, B = 0;
    for (uint32_t a = 0; a < MARSAGLIA_ATTEMPTS; ++a) {
        const int64_t ka = (int64_t)(st.raw(base + 2 * a) >> 8);
        const int64_t kb = (int64_t)(st.raw(base + 2 * a + 1) >> 8);
        const int64_t ta = 2 * ka + 1 - (int64_t(1) << 24);
        const int64_t tb = 2 * kb + 1 - (int64_t(1) << 24);
        if (ta * ta + tb * tb < (int64_t(1) << 48)) { A = ta; B = tb; break; }
    }
    const T q1 = (T)A * (T)5.9604644775390625e-08, q2 = (T)B * (T)5.9604644775390625e-08;
    const T s = q1 * q1 + q2 * q2;
    const T sq = std::sqrt((T)1 - s);
    n[0] = (T)2 * q1 * sq;
    n[1] = (T)2 * q2 * sq;
    n[2] = (T)1 - (T)2 * s;
    normalize3(n[0], n[1], n[2]);
}

template <typename T>
static P4<T> perturbed_normal(Stream& st, int px, int py, const P4<T>& h, T perturbation,
                              uint32_t angle_slot = SLOT_REF_ANGLE) {
    T vx = (T)px, vy = (T)py, vz = (T)1;
    const T vn = std::sqrt(vx * vx + vy * vy + vz * vz);
    vx = vx / vn; vy = vy / vn; vz = vz / vn;
    const T a1 = (uniform<T>(st.raw(angle_slot)) - (T)0.5) * perturbation;
    const T a2 = (uniform<T>(st.raw(angle_slot + 1)) - (T)0.5) * perturbation;
    const T a3 = (uniform<T>(st.raw(angle_slot + 2)) - (T)0.5) * perturbation;
    const T s1 = std::sin(a1), s2 = std::sin(a2), s3 = std::sin(a3);
    const T c1 = std::cos(a1), c2 = std::cos(a2), c3 = std::cos(a3);
    T R[9];
    R[0] = c2 * c3;
    R[1] = c3 * s1 * s2 - c1 * s3;
    R[2] = s1 * s3 + c1 * c3 * s2;
    R[3] = c2 * s3;
    R[4] = c1 * c3 + s1 * s2 * s3;
    R[5] = c1 * s2 * s3 - c3 * s1;
    R[6] = -s2;
    R[7] = c2 * s1;
    R[8] = c1 * c2;
    P4<T> o;
    o.x = R[0] * h.x + R[1] * h.y + R[2] * h.z;
    o.y = R[3] * h.x + R[4] * h.y + R[5] * h.z;
    o.z = R[6] * h.x + R[7] * h.y + R[8] * h.z;
    o.w = h.w;
    if (o.x * vx + o.y * vy + o.z * vz >= (T)0) { o.x = h.x; o.y = h.y; o.z = h.z; }
    normalize3(o.x, o.y, o.z);
    return o;
}

// Neighbour plane hn (owned by neighbour pixel q) re-expressed as a reference-MI hypothesis at p
// (BuildCrossViewPlaneProposal, AdaptMIPMFrame.cu:380-396, denominator as the reference has it).
template <typename T>
static P4<T> reencode_plane(const P4<T>& hn, int qx, int qy, T du, T dv, int px, int py) {
    const T al = -hn.x / hn.z, be = -hn.y / hn.z;
    const T ga = hn.w - al * (T)qx - be * (T)qy;
    T den = (T)1 + al * du + be * dv;
    if (std::fabs(den) < (T)1e-6) den = (den >= (T)0) ? (T)1e-6 : (T)-1e-6;
    const T aR = al / den, bR = be / den, cR = ga / den;
    T nx = -aR, ny = -bR, nz = (T)1;
    normalize3(nx, ny, nz);
    return P4<T>{nx, ny, nz, aR * (T)px + bR * (T)py + cR};
}

template <typename T>
static bool cross_view_proposal(const Frame<T>& f, const P4<T>* plane_prev, int v, int n, int px,
                                int py, const P4<T>& cur, P4<T>& out) {
    const int W = f.W;
    const T sx = ((T)f.centers[2 * v] - (T)f.centers[2 * n]) * f.inv_base;
    const T sy = ((T)f.centers[2 * v + 1] - (T)f.centers[2 * n + 1]) * f.inv_base;
    const T qfx = (T)px + sx * cur.w, qfy = (T)py + sy * cur.w;
    const T fqx = std::floor(qfx + (T)0.5), fqy = std::floor(qfy + (T)0.5);
    if (!(fqx >= (T)0 && fqx < (T)W && fqy >= (T)0 && fqy < (T)W)) return false;
    const int qx = (int)fqx, qy = (int)fqy;
    const P4<T>& hn = plane_prev[(size_t)n * W * W + (size_t)qy * W + qx];
    const T du = (T)f.centers[2 * v] - (T)f.centers[2 * n];
    const T dv = (T)f.centers[2 * v + 1] - (T)f.centers[2 * n + 1];
    out = reencode_plane(hn, qx, qy, du, dv, px, py);
    return true;
}

// The 8 propagation slots of CheckerboardUpdateFrame (AdaptMIPMFrame.cu:584-795) for pixel
// (x, y) of a W x W MI whose previous costs are `cost`: far slots 1/3/5/7 at +-3 along an axis
// with one extra step to +-4 (farDis = 2), near slots 0/2/4/6 at +-1 with the two diagonal
// alternatives of nerDis = 2, each slot taking its lowest previous cost (first of equals).
// Pixels within 3 of a border copy the near slot into the missing far one (:710-795): dup[a] is
// then the slot copied, else -1.  pos[] are pixel indices y*W + x inside the MI; slots with
// flag[a] false hold the centre.
template <typename T>
static void propagation_candidates(int W, int x, int y, const T* cost, bool flag[8], int pos[8],
                                   int dup[8]) {
    const int H = W, center = y * W + x;
    for (int a = 0; a < 8; ++a) { flag[a] = false; pos[a] = center; dup[a] = -1; }
    auto pick = [&](int& best, T& cm, int cand) {
        if (cost[cand] < cm) { cm = cost[cand]; best = cand; }
    };
    if (y > 2) {
        flag[1] = true; int b = center - 3 * W; T cm = cost[b];
        if (y > 3) pick(b, cm, center - 4 * W);
        pos[1] = b;
    }
    if (y < H - 3) {
        flag[3] = true; int b = center + 3 * W; T cm = cost[b];
        if (y < H - 4) pick(b, cm, center + 4 * W);
        pos[3] = b;
    }
    if (x > 2) {
        flag[5] = true; int b = center - 3; T cm = cost[b];
        if (x > 3) pick(b, cm, center - 4);
        pos[5] = b;
    }
    if (x < W - 3) {
        flag[7] = true; int b = center + 3; T cm = cost[b];
        if (x < W - 4) pick(b, cm, center + 4);
        pos[7] = b;
    }
    if (y > 0) {
        flag[0] = true; const int un = center - W; int b = un; T cm = cost[b];
        for (int i = 0; i < 2; ++i) {
            if (y > 1 + i && x > i) pick(b, cm, un - (1 + i) * W - i);
            if (y > 1 + i && x < W - 1 - i) pick(b, cm, un - (1 + i) * W + i);
        }
        pos[0] = b;
    }
    if (y <= 2 && y > 0) { flag[1] = true; pos[1] = pos[0]; dup[1] = 0; }
    if (y < H - 1) {
        flag[2] = true; const int dn = center + W; int b = dn; T cm = cost[b];
        for (int i = 0; i < 2; ++i) {
            if (y < H - 2 - i && x > i) pick(b, cm, dn + (1 + i) * W - i);
            if (y < H - 2 - i && x < W - 1 - i) pick(b, cm, dn + (1 + i) * W + i);
        }
        pos[2] = b;
    }
    if (y >= H - 3 && y < H - 1) { flag[3] = true; pos[3] = pos[2]; dup[3] = 2; }
    if (x > 0) {
        flag[4] = true; const int ln = center - 1; int b = ln; T cm = cost[b];
        for (int i = 0; i < 2; ++i) {
            if (x > 1 + i && y > i) pick(b, cm, ln - (1 + i) - i * W);
            if (x > 1 + i && y < H - 1 - i) pick(b, cm, ln - (1 + i) + i * W);
        }
        pos[4] = b;
    }
    if (x <= 2 && x > 0) { flag[5] = true; pos[5] = pos[4]; dup[5] = 4; }
    if (x < W - 1) {
        flag[6] = true; const int rn = center + 1; int b = rn; T cm = cost[b];
        for (int i = 0; i < 2; ++i) {
            if (x < W - 2 - i && y > i) pick(b, cm, rn + (1 + i) - i * W);
            if (x < W - 2 - i && y < H - 1 - i) pick(b, cm, rn + (1 + i) + i * W);
        }
        pos[6] = b;
    }
    if (x >= W - 3 && x < W - 1) { flag[7] = true; pos[7] = pos[6]; dup[7] = 6; }
}

// View-selection priors (AdaptMIPMFrame.cu:797-815): over the 4 direct neighbours that exist
// (up, down, left, right; the same tests as near slots 0/2/4/6), +0.9 for each view whose bit is
// set in the neighbour's previous selection, +0.1 otherwise.  sel = the MI's W*W selections.
template <typename T>
static void view_priors(int W, int x, int y, const uint32_t* sel, int cnt, T prior[KMAX]) {
    const int H = W, center = y * W + x;
    for (int j = 0; j < KMAX; ++j) prior[j] = 0;
    const int direct[4] = {y > 0 ? center - W : -1, y < H - 1 ? center + W : -1,
                           x > 0 ? center - 1 : -1, x < W - 1 ? center + 1 : -1};
    for (int ii = 0; ii < 4; ++ii) {
        if (direct[ii] < 0) continue;
        const uint32_t sp = sel[direct[ii]];
        for (int j = 0; j < cnt; ++j) prior[j] += ((sp >> j) & 1u) ? (T)0.9 : (T)0.1;
    }
}

// Sampling probabilities, CDF and the 15 view draws of CheckerboardUpdateFrame
// (AdaptMIPMFrame.cu:817-871, TransformPDFToCDF CudaUtil.cu:491-507): for view i over the slots
// present, count = #(cost < thr) with thr = 0.8 exp(-iter^2/90), tmpw = sum exp(-c^2/0.18) of
// those, count_false = #(cost > 1.2); p_i = tmpw/count if count > 2 and count_false < 3, else
// exp(-thr^2/0.32) if count_false < 3, else 0; times the prior.  Each draw r = u - FLT_EPSILON
// picks the first view whose CDF value exceeds r.  No view drawn: all views, weight 1.  prob /
// cdf (if not null) receive the probabilities and CDF.
template <typename T>
static void view_select(T ca[8][KMAX], const bool flag[8], const T prior[KMAX], int cnt,
                        int iter, Stream& st, T vw[KMAX], uint32_t& tsel, T* prob_out,
                        T* cdf_out) {
    T prob[KMAX];
    const T thr = (T)0.8 * std::exp((T)(iter * iter) / (T)-90.0);
    for (int i = 0; i < cnt; ++i) {
        T count = 0, tmpw = 0;
        int cf = 0;
        for (int j = 0; j < 8; ++j) {
            if (!flag[j]) continue;
            const T c = ca[j][i];
            if (c < thr) { tmpw += std::exp(c * c / (T)-0.18); count += (T)1; }
            if (c > (T)1.2) ++cf;
        }
        T pr = 0;
        if (count > (T)2 && cf < 3) pr = tmpw / count;
        else if (cf < 3) pr = std::exp(thr * thr / (T)-0.32);
        prob[i] = pr * prior[i];
        if (prob_out) prob_out[i] = prob[i];
    }
    T psum = 0;
    for (int i = 0; i < cnt; ++i) psum += prob[i];
    for (int i = 0; i < KMAX; ++i) vw[i] = 0;
    if (psum > (T)0) {
        const T inv = (T)1 / psum;
        T cum = 0;
        for (int i = 0; i < cnt; ++i) { cum += prob[i] * inv; prob[i] = cum; }
        if (cdf_out) for (int i = 0; i < cnt; ++i) cdf_out[i] = prob[i];
        for (uint32_t sdraw = 0; sdraw < VIEW_DRAWS; ++sdraw) {
            const T r = uniform<T>(st.raw(SLOT_VIEW_DRAW + sdraw)) - (T)1.1920928955078125e-07;
            for (int id = 0; id < cnt; ++id)
                if (prob[id] > r) { vw[id] += (T)1; break; }
        }
    } else if (cdf_out) {
        for (int i = 0; i < cnt; ++i) cdf_out[i] = 0;
    }
    tsel = 0;
    T wn = 0;
    for (int i = 0; i < cnt; ++i)
        if (vw[i] > (T)0) { tsel |= 1u << i; wn += vw[i]; }
    if (wn <= (T)0) {
        for (int i = 0; i < cnt; ++i) { vw[i] = (T)1; tsel |= 1u << i; }
    }
}

// The 5 refinement hypotheses of PlaneHypothesisRefinement_Frame (AdaptMIPMFrame.cu:400-471)
// with the candidate set of the ACMP intent (AdaptMIPMUtil.cu:181-183): d_rand = best.w +
// 2(u - 0.5) if best cost < 0.1, else uniform in [depth_min, depth_max]; n_rand a Marsaglia
// normal; n_pert the perturbed normal at 0.02 pi; d_pert uniform in [0.98 d, 1.02 d].
// Candidates: (n_best, d_rand), (n_rand, d), (n_rand, d_rand), (n_pert, d), (n_best, d_pert).
template <typename T>
static void refine_candidates(const PmParams& p, int x, int y, const P4<T>& best, T bcost,
                              Stream& st, P4<T> cand[5]) {
    const T pert = (T)0.02;
    const T dnow = best.w;
    T drand;
    if (bcost < (T)0.1)
        drand = (uniform<T>(st.raw(SLOT_REF_DEPTH_RAND)) - (T)0.5) * (T)2 + dnow;
    else
        drand = uniform<T>(st.raw(SLOT_REF_DEPTH_RAND)) * ((T)p.depth_max - (T)p.depth_min) +
                (T)p.depth_min;
    T nr[3];
    random_normal<T>(st, SLOT_REF_MARSAGLIA, nr);
    const P4<T> np = perturbed_normal<T>(st, x, y, best, pert * (T)M_PI);
    const T lo = ((T)1 - pert) * dnow, hi = ((T)1 + pert) * dnow;
    const T dpert = uniform<T>(st.raw(SLOT_REF_DEPTH_PERT)) * (hi - lo) + lo;
    cand[0] = P4<T>{best.x, best.y, best.z, drand};
    cand[1] = P4<T>{nr[0], nr[1], nr[2], dnow};
    cand[2] = P4<T>{nr[0], nr[1], nr[2], drand};
    cand[3] = P4<T>{np.x, np.y, np.z, dnow};
    cand[4] = P4<T>{best.x, best.y, best.z, dpert};
}

template <typename T> struct State {
    P4<T>* plane; T* cost; P4<T>* disp; uint32_t* sel;
};

template <typename T>
static void init_pixel(const Frame<T>& f, State<T>& s, int v, int x, int y) {
    const int W = f.W;
    const size_t idx = (size_t)v * W * W + (size_t)y * W + x;
    Stream st(f.p.seed, f.p.frame_uid, (uint32_t)idx, 0u);
    P4<T> h;
    const T depth = uniform<T>(st.raw(SLOT_INIT_DEPTH)) * ((T)f.p.depth_max - (T)f.p.depth_min) +
                    (T)f.p.depth_min;
    T n[3];
    random_normal<T>(st, SLOT_INIT_MARSAGLIA, n);
    h.x = n[0]; h.y = n[1]; h.z = n[2]; h.w = depth;
    Win<T> ww;
    window_weights(f, v, x, y, ww);
    T cv[KMAX];
    cost_vector(f, ww, v, x, y, h, cv);
    T mean; int best;
    const unsigned m = topk(cv, f.count(v), f.p.top_k, mean, best);
    s.plane[idx] = h;
    s.cost[idx] = mean;
    s.disp[idx] = best >= 0 ? disp_record(f, v, best, h) : P4<T>{0, 0, 0, 0};
    s.sel[idx] = m;
}

// ACMM extras of one level (AdaptMIPMFrameACMM.cu): warm-start seeds of the level (null when
// the level starts from random planes).
template <typename T> struct Acmm {
    bool warm;
    const P4<T>* plane_init;
    const T* cost_init;
    const uint32_t* sel_init;
};

static int popcount32(uint32_t v) { return __builtin_popcount(v); }

// RecomputeSupportCostDisp_Frame (AdaptMIPMFrameACMM.cu:259-279): cost vector of h, top-k mean
// and mask, disparity record of the lowest-cost view.
template <typename T>
static void support(const Frame<T>& f, const Win<T>& ww, int v, int x, int y, const P4<T>& h,
                    T& cost, uint32_t& sel, P4<T>& disp) {
    T cv[KMAX];
    cost_vector(f, ww, v, x, y, h, cv);
    T mean; int best;
    sel = topk(cv, f.count(v), f.p.top_k, mean, best);
    cost = mean;
    disp = best >= 0 ? disp_record(f, v, best, h) : P4<T>{0, 0, 0, 0};
}

// WarmSeedIsValid_Frame (AdaptMIPMFrameACMM.cu:186-194).
template <typename T>
static bool warm_seed_valid(const Frame<T>& f, const P4<T>& h, T cost, uint32_t sel) {
    return h.w >= (T)f.p.depth_min && h.w <= (T)f.p.depth_max && cost < (T)1 &&
           popcount32(sel) >= 3;
}

template <typename T>
static void update_pixel(const Frame<T>& f, const State<T>& prev, State<T>& next, int v, int x,
                         int y, int iter, int phase, const Acmm<T>* ac = nullptr) {
    const int W = f.W;
    const size_t ppv = (size_t)W * W;
    const int center = y * W + x;
    const size_t idx = (size_t)v * ppv + center;
    const int cnt = f.count(v);
    if (((x + y) & 1) != phase || cnt <= 0) {
        next.plane[idx] = prev.plane[idx];
        next.cost[idx] = prev.cost[idx];
        next.disp[idx] = prev.disp[idx];
        next.sel[idx] = prev.sel[idx];
        return;
    }
    auto pprev = [&](int c) { return prev.plane[(size_t)v * ppv + c]; };

    Win<T> ww;
    window_weights(f, v, x, y, ww);

    T ca[8][KMAX];
    bool flag[8];
    int pos[8], dup[8];
    propagation_candidates(W, x, y, prev.cost + (size_t)v * ppv, flag, pos, dup);
    for (int a = 0; a < 8; ++a) {
        if (!flag[a] || dup[a] >= 0) continue;
        cost_vector(f, ww, v, x, y, pprev(pos[a]), ca[a]);
    }
    for (int a = 0; a < 8; ++a) {
        if (!flag[a]) { for (int b = 0; b < KMAX; ++b) ca[a][b] = (T)2; continue; }
        if (dup[a] >= 0) for (int b = 0; b < KMAX; ++b) ca[a][b] = ca[dup[a]][b];
    }

    T prior[KMAX];
    view_priors(W, x, y, prev.sel + (size_t)v * ppv, cnt, prior);
    Stream st(f.p.seed, f.p.frame_uid, (uint32_t)idx, (uint32_t)(1 + 2 * iter + phase));
    T vw[KMAX];
    uint32_t tsel = 0;
    view_select<T>(ca, flag, prior, cnt, iter, st, vw, tsel, nullptr, nullptr);

    P4<T> best = prev.plane[idx];
    T bcost;
    P4<T> bdisp{0, 0, 0, 0};
    int am;
    {
        T cv[KMAX];
        cost_vector(f, ww, v, x, y, best, cv);
        bcost = weighted_cost(cv, vw, cnt, am);
        if (am >= 0) bdisp = disp_record(f, v, am, best);
    }
    for (int i = 0; i < 8; ++i) {
        if (!flag[i]) continue;
        const T c = weighted_cost(ca[i], vw, cnt, am);
        if (c < bcost) {
            bcost = c;
            best = pprev(pos[i]);
            bdisp = am >= 0 ? disp_record(f, v, am, best) : P4<T>{0, 0, 0, 0};
        }
    }
    // cross-view proposals through every selected view
    for (int k = 0; k < cnt; ++k) {
        if (!((tsel >> k) & 1u)) continue;
        const int n = f.nbr[(size_t)v * f.K + k];
        if (n < 0) continue;
        P4<T> prop;
        if (!cross_view_proposal(f, prev.plane, v, n, x, y, best, prop)) continue;
        T cv[KMAX];
        cost_vector(f, ww, v, x, y, prop, cv);
        const T c = weighted_cost(cv, vw, cnt, am);
        if (c < bcost) {
            bcost = c;
            best = prop;
            bdisp = am >= 0 ? disp_record(f, v, am, best) : P4<T>{0, 0, 0, 0};
        }
    }
    // refinement (ACMP depth x normal candidates)
    {
        P4<T> cand[5];
        refine_candidates(f.p, x, y, best, bcost, st, cand);
        for (int i = 0; i < 5; ++i) {
            T cv[KMAX];
            cost_vector(f, ww, v, x, y, cand[i], cv);
            const T c = weighted_cost(cv, vw, cnt, am);
            if (c < bcost) {
                bcost = c;
                best = cand[i];
                bdisp = am >= 0 ? disp_record(f, v, am, best) : P4<T>{0, 0, 0, 0};
            }
        }
    }
    if (ac) {
        // DetailRestorer_FrameACMM (AdaptMIPMFrameACMM.cu:470-503) against the level's warm seed,
        // then the metadata of the final plane from its own top-k support (:881-898).  The
        // reference's best_sel is the sampled selection throughout the update.
        uint32_t bsel = tsel;
        if (ac->warm) {
            const P4<T> seed = ac->plane_init[idx];
            if (warm_seed_valid(f, seed, ac->cost_init[idx], ac->sel_init[idx])) {
                T scost; uint32_t ssel; P4<T> sdisp;
                support(f, ww, v, x, y, seed, scost, ssel, sdisp);
                const int seed_count = popcount32(ssel), best_count = popcount32(bsel);
                const T disp_gap = std::fabs(best.w - seed.w);
                const T large_gap = (T)0.08 * std::max((T)f.p.base, (T)1e-3);
                const bool clearly_better =
                    (bcost + (T)0.10 < scost) || (best_count >= seed_count + 2);
                const bool ambiguous = bcost > scost - (T)0.03;
                const bool maybe_carrier = disp_gap > large_gap && ambiguous &&
                                           best_count <= seed_count;
                if ((!clearly_better && maybe_carrier) || (scost + (T)0.02 < bcost)) {
                    best = seed; bcost = scost; bdisp = sdisp; bsel = ssel;
                }
            }
        }
        support(f, ww, v, x, y, best, bcost, bsel, bdisp);
        next.plane[idx] = best;
        next.cost[idx] = bcost;
        next.disp[idx] = bdisp;
        next.sel[idx] = bsel;
        return;
    }
    next.plane[idx] = best;
    next.cost[idx] = bcost;
    next.disp[idx] = bdisp;
    next.sel[idx] = tsel;
}

// InitializeFrameFromRandomOrWarm (AdaptMIPMFrameACMM.cu:373-472): the random plane of
// init_pixel, then (warm levels) the seed, a jittered seed (disparity x (1 +- 0.15)) and a seed
// with its normal perturbed by 0.015 pi, each kept if its top-k cost is strictly lower.  The
// reference re-encodes the perturbed seed through its own plane (an identity); skipped.
template <typename T>
static void init_pixel_acmm(const Frame<T>& f, State<T>& s, int v, int x, int y,
                            const Acmm<T>& ac) {
    const int W = f.W;
    const size_t idx = (size_t)v * W * W + (size_t)y * W + x;
    Stream st(f.p.seed, f.p.frame_uid, (uint32_t)idx, 0u);
    P4<T> h;
    const T depth = uniform<T>(st.raw(SLOT_INIT_DEPTH)) * ((T)f.p.depth_max - (T)f.p.depth_min) +
                    (T)f.p.depth_min;
    T n[3];
    random_normal<T>(st, SLOT_INIT_MARSAGLIA, n);
    h.x = n[0]; h.y = n[1]; h.z = n[2]; h.w = depth;
    Win<T> ww;
    window_weights(f, v, x, y, ww);
    P4<T> best = h;
    T bcost; uint32_t bsel; P4<T> bdisp;
    support(f, ww, v, x, y, h, bcost, bsel, bdisp);
    if (ac.warm) {
        const P4<T> warm = ac.plane_init[idx];
        if (warm_seed_valid(f, warm, ac.cost_init[idx], ac.sel_init[idx])) {
            T c; uint32_t m; P4<T> d;
            support(f, ww, v, x, y, warm, c, m, d);
            if (c < bcost) { best = warm; bcost = c; bsel = m; bdisp = d; }
            P4<T> jit = warm;
            const T u = uniform<T>(st.raw(SLOT_WARM_JITTER));
            jit.w = std::min((T)f.p.depth_max, std::max((T)f.p.depth_min,
                             warm.w * ((T)1 + (u - (T)0.5) * (T)2 * (T)0.15)));
            support(f, ww, v, x, y, jit, c, m, d);
            if (c < bcost) { best = jit; bcost = c; bsel = m; bdisp = d; }
            const P4<T> np = perturbed_normal<T>(st, x, y, warm, (T)0.015 * (T)M_PI,
                                                 SLOT_WARM_ANGLE);
            const P4<T> ref{np.x, np.y, np.z, warm.w};
            support(f, ww, v, x, y, ref, c, m, d);
            if (c < bcost) { best = ref; bcost = c; bsel = m; bdisp = d; }
        }
    }
    s.plane[idx] = best;
    s.cost[idx] = bcost;
    s.disp[idx] = bdisp;
    s.sel[idx] = bsel;
}

template <typename T>
static void build_frame(Frame<T>& f, const uint8_t* tiles_u8, const double* blur,
                        const float* centers, const int* nbr, const int* cnt, int V,
                        const PmParams& p) {
    f.V = V; f.W = p.crop; f.K = p.max_neighbors;
    f.centers = centers; f.nbr = nbr; f.cnt = cnt; f.p = p;
    f.inv_base = (T)1 / std::max((T)p.base, (T)1e-6);
    const size_t n = (size_t)V * p.crop * p.crop;
    f.img.resize(n);
    for (size_t i = 0; i < n; ++i) f.img[i] = (T)tiles_u8[i] / (T)255;
    if (p.blur_lambda > 0.0f && blur) {  // blur values as the NCC reads them (score / 255)
        f.blur.resize(n);
        for (size_t i = 0; i < n; ++i) f.blur[i] = (T)blur[i];
    }
}

// Stage lines on stderr ("init done", "iteration i phase p done") from solve, for callers that
// run a long solve under a watchdog of output silence.
static int g_progress = 0;

template <typename T>
static void solve(const uint8_t* tiles_u8, const double* blur, const float* centers,
                  const int* nbr, const int* cnt, int V, const PmParams& p, T* plane_out,
                  T* cost_out, T* disp_out, uint32_t* sel_out, int nthreads) {
    Frame<T> f;
    build_frame(f, tiles_u8, blur, centers, nbr, cnt, V, p);
    const size_t n = (size_t)V * p.crop * p.crop;
    std::vector<P4<T>> pa(n), pb(n), da(n), db(n);
    std::vector<T> ca(n), cb(n);
    std::vector<uint32_t> sa(n), sb(n);
    State<T> A{pa.data(), ca.data(), da.data(), sa.data()};
    State<T> B{pb.data(), cb.data(), db.data(), sb.data()};
    const int W = p.crop;
    if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (int v = 0; v < V; ++v)
        for (int y = 0; y < W; ++y)
            for (int x = 0; x < W; ++x) init_pixel(f, A, v, x, y);
    if (g_progress) fprintf(stderr, "oracle: init done (%d MIs)\n", V);
    for (int it = 0; it < p.iterations; ++it) {
        for (int ph = 0; ph < 2; ++ph) {
            const State<T>& src = ph == 0 ? A : B;
            State<T>& dst = ph == 0 ? B : A;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
            for (int v = 0; v < V; ++v)
                for (int y = 0; y < W; ++y)
                    for (int x = 0; x < W; ++x) update_pixel(f, src, dst, v, x, y, it, ph);
            if (g_progress) fprintf(stderr, "oracle: iteration %d phase %d done\n", it, ph);
        }
    }
    for (size_t i = 0; i < n; ++i) {
        plane_out[4 * i] = pa[i].x; plane_out[4 * i + 1] = pa[i].y;
        plane_out[4 * i + 2] = pa[i].z; plane_out[4 * i + 3] = pa[i].w;
        cost_out[i] = ca[i];
        disp_out[4 * i] = da[i].x; disp_out[4 * i + 1] = da[i].y;
        disp_out[4 * i + 2] = da[i].z; disp_out[4 * i + 3] = da[i].w;
        sel_out[i] = sa[i];
    }
}

// ------------------------------------------------------------------------------------------
// ACMM multi-scale variant (AdaptMIPMFrameACMM.cpp:155-665, AdaptMIPMFrameACMM.cu:373-991)
// ------------------------------------------------------------------------------------------
// Area-average resize of one W x W tile to W0 x W0 (cv::resize INTER_AREA as OpenCV documents
// it: every output pixel is the mean of the source area it covers, partial pixels weighted by
// the covered fraction).
template <typename T>
static void area_resize(const T* src, int W, int W0, T* dst) {
    const T S = (T)W / (T)W0;
    std::vector<T> wt((size_t)W0 * W, (T)0);
    for (int d = 0; d < W0; ++d) {
        const T a = (T)d * S, b = std::min((T)(d + 1) * S, (T)W);
        const T cell = b - a;
        for (int sx = 0; sx < W; ++sx) {
            const T lo = std::max(a, (T)sx), hi = std::min(b, (T)(sx + 1));
            if (hi > lo) wt[(size_t)d * W + sx] = (hi - lo) / cell;
        }
    }
    for (int dy = 0; dy < W0; ++dy)
        for (int dx = 0; dx < W0; ++dx) {
            T acc = 0;
            for (int sy = 0; sy < W; ++sy) {
                const T wy = wt[(size_t)dy * W + sy];
                if (wy == (T)0) continue;
                for (int sx = 0; sx < W; ++sx) {
                    const T wx = wt[(size_t)dx * W + sx];
                    if (wx != (T)0) acc += wy * wx * src[(size_t)sy * W + sx];
                }
            }
            dst[(size_t)dy * W0 + dx] = acc;
        }
}

// Level schedule of AdaptMIPMFrameACMM::BuildBlurAlignedPyramids / InitGPUParamsFromCPUParams
// (AdaptMIPMFrameACMM.cpp:155-200, 300-345): two levels when the crop is at least 24 px,
// otherwise one; per level the patch, iteration count and top-k of the reference, Base and the
// disparity range scaled with the level.
struct AcmmLevel { int width; float scale; int radius, iterations, top_k; };

static int acmm_levels(int W, AcmmLevel out[2]) {
    int max_levels = 1, cur = W;
    while (max_levels < 4 && cur >= 24) {
        cur /= 2;
        if (cur < 12) break;
        ++max_levels;
    }
    const int L = std::max(1, std::min(2, max_levels));
    for (int lid = 0; lid < L; ++lid) {
        const int coarse = L - 1 - lid;
        const float scale = 1.0f / std::pow(2.0f, (float)coarse);
        AcmmLevel& l = out[lid];
        l.scale = scale;
        l.width = std::max(8, (int)std::round((float)W * scale));
        if (L <= 1) { l.iterations = 5; l.radius = 2; }
        else if (lid == 0) { l.iterations = 3; l.radius = 3; }
        else { l.iterations = 6; l.radius = 2; }
        l.top_k = lid == 0 ? 6 : 5;
    }
    return L;
}

template <typename T>
static void solve_level(Frame<T>& f, const Acmm<T>& ac, std::vector<P4<T>>& pa,
                        std::vector<T>& ca, std::vector<P4<T>>& da, std::vector<uint32_t>& sa,
                        int nthreads) {
    const int V = f.V, W = f.W;
    const size_t n = (size_t)V * W * W;
    pa.assign(n, P4<T>{0, 0, 1, 0}); ca.assign(n, (T)2); da.assign(n, P4<T>{0, 0, 0, 0});
    sa.assign(n, 0u);
    std::vector<P4<T>> pb(n), db(n);
    std::vector<T> cb(n);
    std::vector<uint32_t> sb(n);
    State<T> A{pa.data(), ca.data(), da.data(), sa.data()};
    State<T> B{pb.data(), cb.data(), db.data(), sb.data()};
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (int v = 0; v < V; ++v)
        for (int y = 0; y < W; ++y)
            for (int x = 0; x < W; ++x) init_pixel_acmm(f, A, v, x, y, ac);
    for (int it = 0; it < f.p.iterations; ++it)
        for (int ph = 0; ph < 2; ++ph) {
            const State<T>& src = ph == 0 ? A : B;
            State<T>& dst = ph == 0 ? B : A;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
            for (int v = 0; v < V; ++v)
                for (int y = 0; y < W; ++y)
                    for (int x = 0; x < W; ++x) update_pixel(f, src, dst, v, x, y, it, ph, &ac);
        }
    // PostRefineMetadataFrame (AdaptMIPMFrameACMM.cu:903-929)
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (int v = 0; v < V; ++v)
        for (int y = 0; y < W; ++y)
            for (int x = 0; x < W; ++x) {
                const size_t i = (size_t)v * W * W + (size_t)y * W + x;
                Win<T> ww;
                window_weights(f, v, x, y, ww);
                support(f, ww, v, x, y, pa[i], ca[i], sa[i], da[i]);
            }
}

// blur: V*crop*crop blur values as the NCC reads them (score / 255) when p.blur_lambda > 0, else
// null; a coarse level reads them area-resized like the intensities (BuildBlurAlignedPyramids,
// AdaptMIPMFrameACMM.cpp:338-350; its Gaussian alignment, :295-302, is switched off).
template <typename T>
static void solve_acmm(const uint8_t* tiles_u8, const double* blur, const float* centers,
                       const int* nbr, const int* cnt, int V, const PmParams& p, T* plane_out,
                       T* cost_out, T* disp_out, uint32_t* sel_out, int nthreads) {
    if (nthreads <= 0) nthreads = 1;
    AcmmLevel lv[2];
    const int L = acmm_levels(p.crop, lv);
    const int W = p.crop;
    std::vector<T> full((size_t)V * W * W);
    for (size_t i = 0; i < full.size(); ++i) full[i] = (T)tiles_u8[i] / (T)255;
    const bool with_blur = p.blur_lambda > 0.0f && blur;
    std::vector<T> full_blur(with_blur ? full.size() : 0);
    for (size_t i = 0; i < full_blur.size(); ++i) full_blur[i] = (T)blur[i];
    std::vector<P4<T>> pa, da, pinit;
    std::vector<T> ca, cinit;
    std::vector<uint32_t> sa, sinit;
    int prev_w = 0;
    float prev_scale = 1.0f;
    for (int lid = 0; lid < L; ++lid) {
        const AcmmLevel& l = lv[lid];
        PmParams q = p;
        q.crop = l.width;
        q.base = p.base * l.scale;
        q.radius = l.radius;
        q.iterations = l.iterations;
        q.top_k = l.top_k;
        q.depth_min = p.depth_min * l.scale;
        q.depth_max = p.depth_max * l.scale;
        std::vector<float> lc(2 * (size_t)V);
        for (size_t i = 0; i < lc.size(); ++i) lc[i] = centers[i] * l.scale;
        Frame<T> f;
        f.V = V; f.W = l.width; f.K = p.max_neighbors;
        f.centers = lc.data(); f.nbr = nbr; f.cnt = cnt; f.p = q;
        f.inv_base = (T)1 / std::max((T)q.base, (T)1e-6);
        const size_t n = (size_t)V * l.width * l.width;
        f.img.resize(n);
        if (l.width == W) {
            f.img = full;
            f.blur = full_blur;
        } else {
            for (int v = 0; v < V; ++v)
                area_resize(full.data() + (size_t)v * W * W, W, l.width,
                            f.img.data() + (size_t)v * l.width * l.width);
            if (with_blur) {
                f.blur.resize(n);
                for (int v = 0; v < V; ++v)
                    area_resize(full_blur.data() + (size_t)v * W * W, W, l.width,
                                f.blur.data() + (size_t)v * l.width * l.width);
            }
        }
        Acmm<T> ac{false, nullptr, nullptr, nullptr};
        if (lid > 0) {
            // PrepareWarmStartFromPreviousLevel (AdaptMIPMFrameACMM.cpp:496-570)
            const int W0 = prev_w, W1 = l.width;
            pinit.assign(n, P4<T>{0, 0, 1, (T)-1});
            cinit.assign(n, (T)2);
            sinit.assign(n, 0u);
            const float sx = (float)W0 / (float)std::max(1, W1);
            const float dscale = l.scale / std::max(prev_scale, 1e-6f);
            const float prev_base = p.base * prev_scale;
            const T thresh = (T)std::max(0.03f, 0.12f * prev_base);
            for (int v = 0; v < V; ++v)
                for (int y = 0; y < W1; ++y) {
                    const int py = std::min(W0 - 1, (int)std::floor((float)y * sx));
                    for (int x = 0; x < W1; ++x) {
                        const int px = std::min(W0 - 1, (int)std::floor((float)x * sx));
                        const size_t pi = (size_t)v * W0 * W0 + (size_t)py * W0 + px;
                        T mean = 0, sq = 0;
                        for (int oy = -1; oy <= 1; ++oy) {
                            const int yy = std::min(W0 - 1, std::max(0, py + oy));
                            for (int ox = -1; ox <= 1; ++ox) {
                                const int xx = std::min(W0 - 1, std::max(0, px + ox));
                                const T d = pa[(size_t)v * W0 * W0 + (size_t)yy * W0 + xx].w;
                                mean += d;
                                sq += d * d;
                            }
                        }
                        mean /= (T)9;
                        const T var = std::max((T)0, sq / (T)9 - mean * mean);
                        const bool reliable = ca[pi] < (T)1 && popcount32(sa[pi]) >= 3 &&
                                              std::sqrt(var) <= thresh;
                        if (!reliable) continue;
                        const size_t ci = (size_t)v * W1 * W1 + (size_t)y * W1 + x;
                        P4<T> h = pa[pi];
                        h.w *= (T)dscale;
                        pinit[ci] = h;
                        cinit[ci] = ca[pi];
                        sinit[ci] = sa[pi];
                    }
                }
            ac = Acmm<T>{true, pinit.data(), cinit.data(), sinit.data()};
        }
        solve_level(f, ac, pa, ca, da, sa, nthreads);
        prev_w = l.width;
        prev_scale = l.scale;
    }
    const size_t n = pa.size();
    for (size_t i = 0; i < n; ++i) {
        plane_out[4 * i] = pa[i].x; plane_out[4 * i + 1] = pa[i].y;
        plane_out[4 * i + 2] = pa[i].z; plane_out[4 * i + 3] = pa[i].w;
        cost_out[i] = ca[i];
        disp_out[4 * i] = da[i].x; disp_out[4 * i + 1] = da[i].y;
        disp_out[4 * i + 2] = da[i].z; disp_out[4 * i + 3] = da[i].w;
        sel_out[i] = sa[i];
    }
}

template <typename T>
static void eval_costs(const uint8_t* tiles_u8, const double* blur, const float* centers,
                       const int* nbr, const int* cnt, int V, const PmParams& p,
                       const int32_t* hyp, int n, T* out, int nthreads,
                       double* margins = nullptr) {
    Frame<T> f;
    build_frame(f, tiles_u8, blur, centers, nbr, cnt, V, p);
    const int K = p.max_neighbors, W = p.crop;
    if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
    for (int i = 0; i < n; ++i) {
        const int32_t* r = hyp + 6 * (size_t)i;
        const int v = r[0], pix = r[1];
        float fl[4];
        std::memcpy(fl, r + 2, sizeof(fl));
        const P4<T> h{(T)fl[0], (T)fl[1], (T)fl[2], (T)fl[3]};
        Win<T> ww;
        const int x = pix % W, y = pix / W;
        window_weights(f, v, x, y, ww);
        if (margins) {
            // per view: cost, then the margins of NccMargins (views past the count: 2)
            for (int k = 0; k < K; ++k) {
                NccMargins mg{INFINITY, INFINITY, 0};
                const T c = k < f.count(v) ? ncc(f, ww, v, k, x, y, h, &mg) : (T)2;
                out[(size_t)i * K + k] = c;
                margins[3 * ((size_t)i * K + k)] = mg.geo;
                margins[3 * ((size_t)i * K + k) + 1] = mg.gate;
                margins[3 * ((size_t)i * K + k) + 2] = mg.ferr;
            }
            continue;
        }
        T cv[KMAX];
        cost_vector(f, ww, v, x, y, h, cv);
        for (int k = 0; k < K; ++k) out[(size_t)i * K + k] = cv[k];
    }
}

// ------------------------------------------------------------------------------------------
// Layout (brute force)
// ------------------------------------------------------------------------------------------
static int resolve_crop(const LayoutParams& lp) {
    return lp.crop > 0 ? lp.crop : (int)(lp.pitch * 0.707106781186f);
}

// Ideal hex centre snapped so that the crop origin c - (crop-1)/2 is an integer.
static void lattice_center(float pitch, int crop, int i, int j, float& cx, float& cy) {
    const double p = (double)pitch, h = (crop - 1) * 0.5;
    const double ix = p * ((double)i + 0.5 * (double)(1 + (j & 1)));
    const double iy = p * (0.5 + 0.8660254037844386 * (double)j);
    cx = (float)(std::floor(ix - h + 0.5) + h);
    cy = (float)(std::floor(iy - h + 0.5) + h);
}

static int layout(const LayoutParams& lp, int cap, float* centers, int* origins, int* tilekeys,
                  int* lens_type, int* nbr, int* cnt) {
    const int crop = resolve_crop(lp);
    const double rs = (double)lp.pitch * 0.8660254037844386;
    const int cols = (int)std::ceil(lp.raw_w / (double)lp.pitch) + 1;
    const int rows = (int)std::ceil(lp.raw_h / rs) + 1;
    struct MI { int i, j; float cx, cy; int ox, oy; };
    std::vector<MI> mis;
    const double half = (crop - 1) * 0.5;
    for (int j = 0; j < rows; ++j)
        for (int i = 0; i < cols; ++i) {
            float cx, cy;
            lattice_center(lp.pitch, crop, i, j, cx, cy);
            const double lx = (double)cx - half, ly = (double)cy - half;
            if (lx < 0 || ly < 0) continue;
            const int ox = (int)std::floor(lx), oy = (int)std::floor(ly);
            if (ox + crop > lp.raw_w || oy + crop > lp.raw_h) continue;
            mis.push_back(MI{i, j, cx, cy, ox, oy});
        }
    const int V = (int)mis.size();
    if (cap < V) return V;
    const int K = lp.max_neighbors;
    for (int v = 0; v < V; ++v) {
        const MI& m = mis[v];
        if (centers) { centers[2 * v] = m.cx; centers[2 * v + 1] = m.cy; }
        if (origins) { origins[2 * v] = m.ox; origins[2 * v + 1] = m.oy; }
        if (tilekeys) { tilekeys[2 * v] = m.i; tilekeys[2 * v + 1] = m.j; }
        const int q = m.i - (m.j - (m.j & 1)) / 2;
        const int t = lp.lens_types == 3 ? (((q - m.j) % 3) + 3) % 3 : 0;
        if (lens_type) lens_type[v] = t;
        if (!nbr && !cnt) continue;
        struct Cand { int d4; double ang; int id; };
        std::vector<Cand> cs;
        int capn = lp.lens_types == 3 ? lp.neighbors_per_type[t] : K;
        capn = std::min(capn, K);
        // distance class of the capn-th neighbour of an interior MI: 6 at 4d^2=4 (ring 1),
        // 6 at 4d^2=12 and 6 at 4d^2=16 (ring 2)
        const int d4max = capn <= 6 ? 4 : (capn <= 12 ? 12 : 16);
        for (int u = 0; u < V; ++u) {
            if (u == v) continue;
            const int dx2 = 2 * (mis[u].i - m.i) + ((mis[u].j & 1) - (m.j & 1));
            const int dy = mis[u].j - m.j;
            const int d4 = dx2 * dx2 + 3 * dy * dy;
            if (d4 > d4max) continue;
            double a = std::atan2(std::sqrt(3.0) * dy, (double)dx2);
            if (a < 0) a += 2.0 * M_PI;
            cs.push_back(Cand{d4, a, u});
        }
        std::sort(cs.begin(), cs.end(), [](const Cand& a, const Cand& b) {
            return a.d4 != b.d4 ? a.d4 < b.d4 : a.ang < b.ang;
        });
        const int c = std::min((int)cs.size(), capn);
        if (cnt) cnt[v] = c;
        if (nbr)
            for (int k = 0; k < K; ++k) nbr[(size_t)v * K + k] = k < c ? cs[k].id : -1;
    }
    return V;
}

// FixedPosition mode-1 key offsets of key (X, Y) in a u x v lens lattice, in collection
// order: the reference's if/else chain (SelectNeighbors.cpp:1062-1286) branch for branch,
// including its top-row test, which holds for every X on row 0.
static int fixed_offsets(int X, int Y, int u, int v, int out[12][2]) {
    int n = 0;
    auto add = [&](int dx, int dy) { out[n][0] = dx; out[n][1] = dy; ++n; };
    if (X > 1 && X < u - 2 && Y > 0 && Y < v - 1) {
        add(1, 0); add(-1, 0); add(2, 0); add(0, -1); add(-2, -1); add(0, 1);
        add(-2, 0); add(-1, -1); add(1, -1); add(-1, 1); add(1, 1); add(2, 1);
    } else if (X == 0 && Y == 0) {
        add(1, 0); add(2, 0); add(0, 1); add(1, 1); add(2, 2); add(0, 2);
    } else if ((X == 1 && Y == 0) || ((X != u - 2 || X != u - 1) && Y == 0)) {
        add(1, 0); add(-1, 0); add(2, 0); add(0, 1); add(-1, 1); add(1, 1);
    } else if (X == 0 && Y == v - 1) {
        add(1, 0); add(2, 0); add(0, -1); add(1, -1); add(2, -2); add(0, -2);
    } else if (X == 1 && Y == v - 1) {
        add(1, 0); add(-1, 0); add(2, 0); add(0, -1); add(-1, -1); add(1, -1);
    } else if (X == u - 1 && Y == v - 1) {
        add(-1, 0); add(0, -1); add(-2, -1); add(-2, 0); add(0, -2); add(-1, -1);
    } else if (X == u - 2 && Y == v - 1) {
        add(1, 0); add(-1, 0); add(0, -1); add(-2, -1); add(-1, -1); add(1, -1);
    } else if (X == u - 1 && Y == 0) {
        add(-1, 0); add(0, 1); add(-2, 2); add(-2, 0); add(0, 2); add(-1, 1);
    } else if (X == u - 2 && Y == 0) {
        add(1, 0); add(-1, 0); add(0, 1); add(-2, 0); add(-1, 1); add(1, 1);
    } else if (X == 0) {
        add(1, 0); add(2, 0); add(0, -1); add(0, 1); add(1, -1); add(1, 1);
    } else if (X == 1) {
        add(1, 0); add(-1, 0); add(2, 0); add(0, -1); add(0, 1); add(-1, -1);
    } else if (Y == v - 1) {
        add(1, 0); add(-1, 0); add(2, 0); add(0, -1); add(-2, -1); add(1, -1);
    } else if (X == u - 1) {
        add(-1, 0); add(0, -1); add(-2, -1); add(0, 1); add(-1, -1); add(-1, 1);
    } else if (X == u - 2) {
        add(1, 0); add(-1, 0); add(0, -1); add(-2, -1); add(0, 1); add(1, 1);
    }
    return n;
}

// FixedPosition table of a layout: kept neighbours sorted by centre distance (insertion sort,
// so equal distances keep the collection order), cap K; garbage = key column/row 0 or none.
static int fixed_neighbors(const LayoutParams& lp, int* nbr, int* cnt, int* garbage) {
    const int V = layout(lp, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    std::vector<float> cen(2 * (size_t)V);
    std::vector<int> keys(2 * (size_t)V);
    layout(lp, V, cen.data(), nullptr, keys.data(), nullptr, nullptr, nullptr);
    const double rs = (double)lp.pitch * 0.8660254037844386;
    const int u = (int)std::ceil(lp.raw_w / (double)lp.pitch) + 1;
    const int v = (int)std::ceil(lp.raw_h / rs) + 1;
    std::vector<int> cell((size_t)u * v, -1);
    for (int m = 0; m < V; ++m) cell[(size_t)keys[2 * m + 1] * u + keys[2 * m]] = m;
    const int K = lp.max_neighbors;
    for (int m = 0; m < V; ++m) {
        const int X = keys[2 * m], Y = keys[2 * m + 1];
        int off[12][2];
        const int n = fixed_offsets(X, Y, u, v, off);
        std::vector<std::pair<float, int>> list;  // (distance, MI), insertion-sorted
        for (int e = 0; e < n; ++e) {
            const int nx = X + off[e][0], ny = Y + off[e][1];
            if (nx < 0 || ny < 0 || nx >= u || ny >= v) continue;
            const int id = cell[(size_t)ny * u + nx];
            if (id < 0) continue;
            const float dx = cen[2 * id] - cen[2 * m], dy = cen[2 * id + 1] - cen[2 * m + 1];
            const float len = (float)std::sqrt(std::pow(dy, 2) + std::pow(dx, 2));
            size_t at = list.size();
            while (at > 0 && list[at - 1].first > len) --at;
            list.insert(list.begin() + (long)at, std::make_pair(len, id));
        }
        const int c = std::min((int)list.size(), K);
        if (cnt) cnt[m] = c;
        if (nbr)
            for (int k = 0; k < K; ++k) nbr[(size_t)m * K + k] = k < c ? list[k].second : -1;
        if (garbage) garbage[m] = (list.empty() || X < 1 || Y < 1) ? 1 : 0;
    }
    return V;
}

// ------------------------------------------------------------------------------------------
// Fused CheckModule filter, FP32 twin (bit-exact contract with the GPU pass)
// ------------------------------------------------------------------------------------------
struct FilterCtx {
    int V, W, K;
    const float* gray; const float* centers; const int* nbr; const int* cnt;
    float base;
    const float* plane; const float* cost;
    FilterParams fp;
    float G(int v, int x, int y) const { return gray[(size_t)v * W * W + (size_t)y * W + x]; }
    float D(int v, int i) const { return plane[4 * ((size_t)v * W * W + i) + 3]; }
    float C(int v, int i) const { return cost[(size_t)v * W * W + i]; }
};

static float f_nan() { return std::nanf(""); }

static float bilin_disp(const FilterCtx& c, int n, float x, float y) {
    const int W = c.W;
    if (x < 0 || y < 0 || x > W - 1 || y > W - 1) return f_nan();
    const int x0 = (int)std::floor(x), y0 = (int)std::floor(y);
    const int x1 = std::min(x0 + 1, W - 1), y1 = std::min(y0 + 1, W - 1);
    const float ax = x - (float)x0, ay = y - (float)y0;
    auto dat = [&](int xi, int yi) {
        const float d = c.D(n, yi * W + xi);
        if (!std::isfinite(d) || d <= c.fp.min_valid_disp) return f_nan();
        return d;
    };
    const float d00 = dat(x0, y0), d01 = dat(x0, y1), d10 = dat(x1, y0), d11 = dat(x1, y1);
    if (!std::isfinite(d00) || !std::isfinite(d01) || !std::isfinite(d10) || !std::isfinite(d11))
        return f_nan();
    return (1.0f - ax) * (1.0f - ay) * d00 + (1.0f - ax) * ay * d01 + ax * (1.0f - ay) * d10 +
           ax * ay * d11;
}

static float bilin_cost(const FilterCtx& c, int n, float x, float y) {
    const int W = c.W;
    if (x < 0 || y < 0 || x > W - 1 || y > W - 1) return f_nan();
    const int x0 = (int)std::floor(x), y0 = (int)std::floor(y);
    const int x1 = std::min(x0 + 1, W - 1), y1 = std::min(y0 + 1, W - 1);
    const float ax = x - (float)x0, ay = y - (float)y0;
    const float c00 = c.C(n, y0 * W + x0), c01 = c.C(n, y1 * W + x0);
    const float c10 = c.C(n, y0 * W + x1), c11 = c.C(n, y1 * W + x1);
    if (!std::isfinite(c00) || !std::isfinite(c01) || !std::isfinite(c10) || !std::isfinite(c11))
        return f_nan();
    return (1.0f - ax) * (1.0f - ay) * c00 + (1.0f - ax) * ay * c01 + ax * (1.0f - ay) * c10 +
           ax * ay * c11;
}

static bool bilin_gray(const FilterCtx& c, int v, float x, float y, float& out) {
    const int W = c.W;
    if (x < 0 || y < 0 || x > W - 1 || y > W - 1) return false;
    const int x0 = (int)std::floor(x), y0 = (int)std::floor(y);
    const int x1 = std::min(x0 + 1, W - 1), y1 = std::min(y0 + 1, W - 1);
    const float ax = x - (float)x0, ay = y - (float)y0;
    const float v00 = c.G(v, x0, y0), v01 = c.G(v, x0, y1), v10 = c.G(v, x1, y0),
                v11 = c.G(v, x1, y1);
    out = (1.0f - ax) * (1.0f - ay) * v00 + (1.0f - ax) * ay * v01 + ax * (1.0f - ay) * v10 +
          ax * ay * v11;
    return true;
}

static bool mean_std3(const FilterCtx& c, int v, float fx, float fy, float& mean, float& sd) {
    const int cx = (int)std::round(fx), cy = (int)std::round(fy);
    if (cx < 1 || cy < 1 || cx >= c.W - 1 || cy >= c.W - 1) return false;
    double s1 = 0, s2 = 0;
    for (int yy = cy - 1; yy <= cy + 1; ++yy)
        for (int xx = cx - 1; xx <= cx + 1; ++xx) {
            const double g = (double)c.G(v, xx, yy);
            s1 += g;
            s2 += g * g;
        }
    const double scale = 1.0 / 9.0;
    const double m = s1 * scale;
    const double var = std::max(s2 * scale - m * m, 0.0);
    mean = (float)m;
    sd = (float)std::sqrt(var);
    return std::isfinite(mean) && std::isfinite(sd) && sd > 1.0e-3f;
}

static float photo_residual(const FilterCtx& c, int v, int n, float px, float py, float qx,
                            float qy) {
    float Ir = 0, In = 0;
    if (!bilin_gray(c, v, px, py, Ir)) return f_nan();
    if (!bilin_gray(c, n, qx, qy, In)) return f_nan();
    float mr, sr, mn, sn;
    if (mean_std3(c, v, px, py, mr, sr) && mean_std3(c, n, qx, qy, mn, sn)) {
        const float a = (Ir - mr) / sr;
        const float b = (In - mn) / sn;
        return std::fabs(a - b);
    }
    return std::fabs(In - Ir);
}

static void filter_f32(const FilterCtx& c, const float* disp_in, const uint32_t* sel_in,
                       float* plane_out, float* cost_out, float* disp_out, uint32_t* sel_out,
                       uint8_t* reason_out, int nthreads) {
    const int W = c.W, N = W * W;
    const FilterParams& fp = c.fp;
    if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads)
    for (int v = 0; v < c.V; ++v) {
        std::vector<uint8_t> k1(N, 0), kf(N, 0), reason(N, PR_KEEP);
        const int cntv = std::min(c.cnt[v], c.K);
        for (int y = 0; y < W; ++y)
            for (int x = 0; x < W; ++x) {
                const int i = y * W + x;
                const size_t g = (size_t)v * N + i;
                const float d = c.D(v, i);
                if (!std::isfinite(d) || d <= fp.min_valid_disp) continue;
                const float cost = c.C(v, i);
                if (fp.enable_cost_filter) {
                    const bool invalid = !std::isfinite(cost) || cost < 0.0f;
                    const bool high = std::isfinite(cost) && cost > fp.max_cost;
                    if (invalid || high) { reason[i] = PR_COST; continue; }
                }
                int vids[32];
                int nv = 0;
                if (fp.use_selected_views_only) {
                    uint32_t m = sel_in[g];
                    for (int b = 0; m && nv < fp.max_triplet; ++b) {
                        if (m & 1u) vids[nv++] = b;
                        m >>= 1u;
                    }
                } else {
                    const int lim = std::min(cntv, fp.max_triplet);
                    for (int b = 0; b < lim; ++b) vids[nv++] = b;
                }
                if (nv == 0) { reason[i] = PR_NOSUPPORT; continue; }
                const float px = (float)x, py = (float)y;
                int support = 0, good = 0;
                for (int vi = 0; vi < nv; ++vi) {
                    const int b = vids[vi];
                    if (b < 0 || b >= cntv) continue;
                    const int n = c.nbr[(size_t)v * c.K + b];
                    if (n < 0) continue;
                    const float Bx = c.centers[2 * v] - c.centers[2 * n];
                    const float By = c.centers[2 * v + 1] - c.centers[2 * n + 1];
                    const float bl = std::sqrt(Bx * Bx + By * By);
                    if (bl < 1.0e-6f) continue;
                    const float dx = Bx / bl, dy = By / bl;
                    const float scale = bl / c.base;
                    const float qx = px + dx * d * scale;
                    const float qy = py + dy * d * scale;
                    if (qx < 0 || qy < 0 || qx > W - 1 || qy > W - 1) continue;
                    const float dn = bilin_disp(c, n, qx, qy);
                    if (!std::isfinite(dn)) continue;
                    if (fp.enable_cost_filter) {
                        const float nc = bilin_cost(c, n, qx, qy);
                        const bool inv = !std::isfinite(nc) || nc < 0.0f;
                        const bool hi = std::isfinite(nc) && nc > fp.max_cost;
                        if (inv || hi) continue;
                    }
                    ++support;
                    bool pg = true, pp = true;
                    if (fp.enable_cycle_geo_filter) {
                        const float phx = qx - dx * dn * scale;
                        const float phy = qy - dy * dn * scale;
                        const float ex = phx - px, ey = phy - py;
                        const float ge = std::sqrt(ex * ex + ey * ey);
                        pg = std::isfinite(ge) && ge <= fp.max_geo_err_px;
                    }
                    if (fp.enable_cycle_photo_filter) {
                        const float pe = photo_residual(c, v, n, px, py, qx, qy);
                        pp = std::isfinite(pe) && pe <= fp.max_photo_err_u;
                    }
                    if (pg && pp) ++good;
                }
                if (support < fp.min_good_neighbors) { reason[i] = PR_NOSUPPORT; continue; }
                if (good < fp.min_good_neighbors) { reason[i] = PR_CYCLE; continue; }
                k1[i] = 1; kf[i] = 1; reason[i] = PR_KEEP;
            }
        if (fp.enable_spike_filter) {
            for (int y = 0; y < W; ++y)
                for (int x = 0; x < W; ++x) {
                    const int i = y * W + x;
                    if (!k1[i]) continue;
                    float vals[8];
                    int nvals = 0;
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            if (dx == 0 && dy == 0) continue;
                            const int xn = x + dx, yn = y + dy;
                            if (xn < 0 || yn < 0 || xn >= W || yn >= W) continue;
                            const int ni = yn * W + xn;
                            if (!k1[ni]) continue;
                            const float dv = c.D(v, ni);
                            if (!std::isfinite(dv) || dv <= fp.min_valid_disp) continue;
                            vals[nvals++] = dv;
                        }
                    if (nvals < fp.spike_min_neighbors) continue;
                    std::sort(vals, vals + nvals);
                    const float med = vals[nvals / 2];
                    const float cur = c.D(v, i);
                    if (!std::isfinite(cur)) continue;
                    if (std::fabs(cur - med) > fp.spike_abs_diff) {
                        kf[i] = 0;
                        reason[i] = PR_SPIKE;
                    }
                }
        }
        for (int i = 0; i < N; ++i) {
            const size_t g = (size_t)v * N + i;
            const float d = c.D(v, i);
            const bool drop = std::isfinite(d) && d > fp.min_valid_disp && !kf[i];
            for (int q = 0; q < 4; ++q) {
                plane_out[4 * g + q] = c.plane[4 * g + q];
                disp_out[4 * g + q] = disp_in[4 * g + q];
            }
            cost_out[g] = c.cost[g];
            sel_out[g] = sel_in[g];
            if (drop) {
                plane_out[4 * g] = 0.0f; plane_out[4 * g + 1] = 0.0f;
                plane_out[4 * g + 2] = 0.0f; plane_out[4 * g + 3] = -1.0f;
                cost_out[g] = 2.0f;
                disp_out[4 * g] = -1.0f; disp_out[4 * g + 1] = -1.0f;
                disp_out[4 * g + 2] = 0.0f; disp_out[4 * g + 3] = 0.0f;
                if (fp.clear_selected_views_when_invalid) sel_out[g] = 0u;
            }
            reason_out[g] = reason[i];
        }
    }
}

// ------------------------------------------------------------------------------------------
// Refocus (FP64 samples; MI inclusion decided with the FP32 expression the product uses)
// ------------------------------------------------------------------------------------------
static void refocus(const uint8_t* tiles_u8, const float* centers, int V, float pitch, int crop,
                    int out_w, int out_h, int S, float vmin, float vmax, double* out,
                    int nthreads) {
    const int W = crop;
    // uniform bucket grid of MI centres (cell = pitch)
    const int gw = (int)std::ceil(out_w / (double)pitch) + 4;
    const int gh = (int)std::ceil(out_h / (double)pitch) + 4;
    std::vector<std::vector<int>> cell((size_t)gw * gh);
    for (int v = 0; v < V; ++v) {
        int cx = (int)std::floor(centers[2 * v] / pitch) + 2;
        int cy = (int)std::floor(centers[2 * v + 1] / pitch) + 2;
        cx = std::min(std::max(cx, 0), gw - 1);
        cy = std::min(std::max(cy, 0), gh - 1);
        cell[(size_t)cy * gw + cx].push_back(v);
    }
    const float halff = (float)(crop - 1) * 0.5f;
    const double half = (crop - 1) * 0.5;
    if (nthreads <= 0) nthreads = 1;
    for (int k = 0; k < S; ++k) {
        const float vk = S > 1 ? vmin + (vmax - vmin) * (float)k / (float)(S - 1) : vmin;
        const double reach = (double)vk * half + 1.0;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
        for (int y = 0; y < out_h; ++y)
            for (int x = 0; x < out_w; ++x) {
                const int gx0 = std::max((int)std::floor((x - reach) / pitch) + 2, 0);
                const int gx1 = std::min((int)std::floor((x + reach) / pitch) + 2, gw - 1);
                const int gy0 = std::max((int)std::floor((y - reach) / pitch) + 2, 0);
                const int gy1 = std::min((int)std::floor((y + reach) / pitch) + 2, gh - 1);
                double acc = 0;
                int nacc = 0;
                for (int gy = gy0; gy <= gy1; ++gy)
                    for (int gx = gx0; gx <= gx1; ++gx)
                        for (int v : cell[(size_t)gy * gw + gx]) {
                            const float cx = centers[2 * v], cy = centers[2 * v + 1];
                            const float ufx = ((float)x - cx) / vk + halff;
                            const float ufy = ((float)y - cy) / vk + halff;
                            if (!(ufx >= 0.0f && ufx <= (float)(W - 1) && ufy >= 0.0f &&
                                  ufy <= (float)(W - 1)))
                                continue;
                            const double ux = ((double)x - cx) / (double)vk + half;
                            const double uy = ((double)y - cy) / (double)vk + half;
                            // clamp the FP64 position to the decided inclusion box
                            const double sx = std::min(std::max(ux, 0.0), (double)(W - 1));
                            const double sy = std::min(std::max(uy, 0.0), (double)(W - 1));
                            const double fx0 = std::floor(sx), fy0 = std::floor(sy);
                            const int x0 = (int)fx0, y0 = (int)fy0;
                            const int x1 = std::min(x0 + 1, W - 1), y1 = std::min(y0 + 1, W - 1);
                            const double fx = sx - fx0, fy = sy - fy0;
                            const uint8_t* t = tiles_u8 + (size_t)v * W * W;
                            const double a = t[y0 * W + x0] / 255.0, b = t[y0 * W + x1] / 255.0;
                            const double c = t[y1 * W + x0] / 255.0, d = t[y1 * W + x1] / 255.0;
                            const double top = a + fx * (b - a), bot = c + fx * (d - c);
                            acc += top + fy * (bot - top);
                            ++nacc;
                        }
                out[((size_t)k * out_h + y) * out_w + x] = nacc ? acc / nacc : 0.0;
            }
    }
}

// ------------------------------------------------------------------------------------------
// V2R behavioural model fit: Householder QR least squares, FP64
// ------------------------------------------------------------------------------------------
static int fit_v2r(const float* aL, const float* v, int n, double c[3]) {
    if (n < 3) return -1;
    std::vector<double> X((size_t)n * 3), y(n);
    for (int i = 0; i < n; ++i) {
        const double a = aL[i], vv = v[i];
        X[(size_t)i * 3] = a * vv; X[(size_t)i * 3 + 1] = vv; X[(size_t)i * 3 + 2] = 1.0;
        y[i] = a;
    }
    for (int k = 0; k < 3; ++k) {
        double nrm = 0;
        for (int i = k; i < n; ++i) nrm += X[(size_t)i * 3 + k] * X[(size_t)i * 3 + k];
        nrm = std::sqrt(nrm);
        if (nrm == 0) return -2;
        const double x0 = X[(size_t)k * 3 + k];
        const double alpha = x0 > 0 ? -nrm : nrm;
        std::vector<double> u(n - k);
        for (int i = k; i < n; ++i) u[i - k] = X[(size_t)i * 3 + k];
        u[0] -= alpha;
        double un = 0;
        for (double e : u) un += e * e;
        if (un == 0) continue;
        for (int j = k; j < 3; ++j) {
            double dot = 0;
            for (int i = k; i < n; ++i) dot += u[i - k] * X[(size_t)i * 3 + j];
            const double s = 2.0 * dot / un;
            for (int i = k; i < n; ++i) X[(size_t)i * 3 + j] -= s * u[i - k];
        }
        double dot = 0;
        for (int i = k; i < n; ++i) dot += u[i - k] * y[i];
        const double s = 2.0 * dot / un;
        for (int i = k; i < n; ++i) y[i] -= s * u[i - k];
    }
    for (int k = 2; k >= 0; --k) {
        double s = y[k];
        for (int j = k + 1; j < 3; ++j) s -= X[(size_t)k * 3 + j] * c[j];
        c[k] = s / X[(size_t)k * 3 + k];
    }
    return 0;
}

// ------------------------------------------------------------------------------------------
// Per-frame products, written as the reference's loops (DepthSolver.cpp:6716-6897 and
// :7832-8000) over MIs in map order with per-pixel sample vectors, with the fixes the product
// documents: plain truncation of the mosaic corner (no tolower), VD median indices clamped to
// n-1, float outputs instead of the uint8 renderings.
// ------------------------------------------------------------------------------------------
static void disparity_mosaic(const float* plane, const float* cost, const float* centers, int V,
                             int crop, float max_cost, int raw_w, int raw_h, float* out,
                             float* range) {
    std::fill(out, out + (size_t)raw_w * raw_h, 0.0f);
    float dmin = INFINITY, dmax = -INFINITY;
    const int WW = crop * crop;
    for (int v = 0; v < V; ++v)
        for (int i = 0; i < WW; ++i) {
            const float d = plane[4 * ((size_t)v * WW + i) + 3];
            if (d <= 0.0f) continue;
            dmin = std::min(dmin, d);
            dmax = std::max(dmax, d);
        }
    for (int v = 0; v < V; ++v) {
        const float ox = (float)(centers[2 * v] - (std::round(crop * 0.5) - 1));
        const float oy = (float)(centers[2 * v + 1] - (std::round(crop * 0.5) - 1));
        for (int y = 0; y < crop; ++y)
            for (int x = 0; x < crop; ++x) {
                const size_t i = (size_t)v * WW + y * crop + x;
                if (cost[i] > max_cost) continue;
                const float d = plane[4 * i + 3];
                if (d <= 0.0f) continue;
                const int X = (int)ox + x, Y = (int)oy + y;
                if (X < 0 || X >= raw_w || Y < 0 || Y >= raw_h) continue;
                out[(size_t)Y * raw_w + X] = d;
            }
    }
    if (range) { range[0] = dmin; range[1] = dmax; }
}

// Colour samples come from the gray crops (gray, one channel) or the colour crops (bgr:
// V x crop^2 x 3 bytes B, G, R, three channels, each sorted and filtered on its own as the
// reference's color0/1/2 lists and the 3-channel medianBlur are); color_out has N x channels
// bytes, interleaved.
static void vd_map(const float* plane, const float* centers, const float* gray, const uint8_t* bgr,
                   int V, int crop, float mla_base, int raw_w, int raw_h, float* vd_out,
                   uint8_t* color_out, float* range) {
    const size_t N = (size_t)raw_w * raw_h;
    const int C = bgr ? 3 : 1;
    std::vector<std::vector<float>> vd(N);
    std::vector<std::vector<int>> col(N * C);
    const int WW = crop * crop;
    for (int v = 0; v < V; ++v)
        for (int y = 0; y < crop; ++y)
            for (int x = 0; x < crop; ++x) {
                const size_t i = (size_t)v * WW + y * crop + x;
                float disp = plane[4 * i + 3];
                if (disp <= 0.0f) disp = 1;
                const float f_v = mla_base / disp;
                const int xm = (int)(x - ((crop - 1) * 0.5 + 1));
                const float Xv = xm * f_v + centers[2 * v];
                const int ym = (int)(y - ((crop - 1) * 0.5 + 1));
                const float Yv = ym * f_v + centers[2 * v + 1];
                const int vy = (int)std::round(Yv), vx = (int)std::round(Xv);
                if (vx > 0 && vx < raw_w && vy > 0 && vy < raw_h) {
                    const size_t q = (size_t)vy * raw_w + vx;
                    vd[q].push_back(f_v);
                    for (int c = 0; c < C; ++c)
                        col[q * C + c].push_back(bgr ? (int)bgr[3 * i + c]
                                                     : (gray ? (int)(uint8_t)gray[i] : 0));
                }
            }
    std::vector<float> vdm(N, 0.0f);
    std::vector<uint8_t> cm(N * C, 0);
    float vmin = INFINITY, vmax = -INFINITY;
    for (size_t p = 0; p < N; ++p) {
        const int n = (int)vd[p].size();
        if (n == 0) continue;
        std::sort(vd[p].begin(), vd[p].end());
        const int i1 = std::min(n / 2 + 1, n - 1), i2 = std::min((n + 1) / 2, n - 1);
        vdm[p] = (vd[p][i1] + vd[p][i2]) / 2;
        for (int c = 0; c < C; ++c) {
            std::vector<int>& cl = col[p * C + c];
            std::sort(cl.begin(), cl.end());
            cm[p * C + c] = (uint8_t)((cl[n / 2] + cl[(n + 1) / 2 - 1]) / 2);
        }
        vmin = std::min(vmin, vdm[p]);
        vmax = std::max(vmax, vdm[p]);
    }
    for (size_t p = 0; p < N; ++p) vd_out[p] = vdm[p];
    if (range) { range[0] = vmin; range[1] = vmax; }
    if (!color_out) return;
    std::vector<uint8_t> img(N * C, 0);
    const long long rows = raw_h, cols = raw_w;
    for (long long y = 0; y < rows; ++y)
        for (long long x = 0; x < cols; ++x) {
            const long long p = y * cols + x;
            if (!vd[p].empty()) {
                for (int c = 0; c < C; ++c) img[p * C + c] = cm[p * C + c];
                continue;
            }
            long long t = 4, k = p;
            while (true) {
                switch (t % 4) {
                    case 0: k = y * cols + x + t / 4; break;
                    case 1: k = y * cols + x - t / 4; break;
                    case 2: k = (y + t / 4) * cols + x - t / 4; break;
                    case 3: k = (y - t / 4) * cols + x; break;
                }
                ++t;
                if (k >= rows * cols || k < 0) break;
                if (!vd[k].empty()) {
                    for (int c = 0; c < C; ++c) img[p * C + c] = cm[k * C + c];
                    break;
                }
            }
        }
    std::vector<uint8_t> tmp(N * C);
    for (int it = 0; it < 10; ++it) {
        for (long long y = 0; y < rows; ++y)
            for (long long x = 0; x < cols; ++x)
                for (int c = 0; c < C; ++c) {
                    int w[25], n = 0;
                    for (int j = -2; j <= 2; ++j)
                        for (int i = -2; i <= 2; ++i) {
                            const long long yy = std::min(std::max(y + j, 0LL), rows - 1);
                            const long long xx = std::min(std::max(x + i, 0LL), cols - 1);
                            w[n++] = img[(yy * cols + xx) * C + c];
                        }
                    std::nth_element(w, w + 12, w + 25);
                    tmp[(y * cols + x) * C + c] = (uint8_t)w[12];
                }
        img.swap(tmp);
    }
    std::copy(img.begin(), img.end(), color_out);
}


// ------------------------------------------------------------------------------------------
// All-in-focus composite: LFRefocus::AIFImageCompositeForMIA (FocalStack/LFRefocus.cpp:183-425)
// with FilterMatchingPoints (:2283-2301), GenerateVirtualImage (:2303-2328) and HoleFilling
// (:2362-2498).  Renderings (MAD clamp, normalisation, gamma, colour maps, binary image) are
// display-only and not restated.

// Five-by-five median passes with replicated borders (cv::medianBlur, ksize 5) on an
// interleaved image of C channels, each channel on its own.
static void median5_passes(std::vector<uint8_t>& img, long long rows, long long cols, int C,
                           int passes) {
    std::vector<uint8_t> tmp(img.size());
    for (int it = 0; it < passes; ++it) {
        for (long long y = 0; y < rows; ++y)
            for (long long x = 0; x < cols; ++x)
                for (int c = 0; c < C; ++c) {
                    int w[25], n = 0;
                    for (int j = -2; j <= 2; ++j)
                        for (int i = -2; i <= 2; ++i) {
                            const long long yy = std::min(std::max(y + j, 0LL), rows - 1);
                            const long long xx = std::min(std::max(x + i, 0LL), cols - 1);
                            w[n++] = img[(yy * cols + xx) * C + c];
                        }
                    std::nth_element(w, w + 12, w + 25);
                    tmp[(y * cols + x) * C + c] = (uint8_t)w[12];
                }
        img.swap(tmp);
    }
}

struct VIPoint {
    int mi, sx, sy;   // source MI (layout index), source pixel column / row
    float vd;         // virtual depth baseline / disparity
    float score;      // FilterMatchingPoints score
};

// stats (8 ints): collected points, chosen pixels, hole-filling passes, fills, largest number of
// candidates at one pixel, the MI filled points are attributed to, pixels with more than 16
// candidates, initial count of pixels with every channel >= 1.
static void aif_composite(const float* plane, const float* centers, const uint8_t* bgr,
                          const uint8_t* rich, const uint8_t* blur, const uint8_t* garbage, int V,
                          int crop, float baseline, float B, float bl0, float f_main, int raw_w,
                          int raw_h, float* vd_out, float* inv_out, float* real_mm_out,
                          uint8_t* aif_out, uint8_t* filled_out, uint8_t* denoise_out,
                          int* stats) {
    const int W = crop, WW = crop * crop;
    const long long rows = raw_h, cols = raw_w, N = rows * cols;
    // Step 1 (:207-282): every pixel with d > 0 maps to round(offset * vd + centre); positions
    // in [0, cols] x [0, rows] are collected (the last row and column past the image are never
    // visited by step 2).  Collection order: MI (key) order, then pixel rows, then columns.
    std::vector<std::vector<VIPoint>> co((size_t)(rows + 1) * (cols + 1));
    VIPoint last{-1, 0, 0, 0.0f, 0.0f};
    long long ncollected = 0;
    for (int v = 0; v < V; ++v) {
        if (garbage && garbage[v]) continue;
        const float cx = centers[2 * v], cy = centers[2 * v + 1];
        for (int y = 0; y < W; ++y)
            for (int x = 0; x < W; ++x) {
                const size_t i = (size_t)v * WW + (size_t)y * W + x;
                const float disp = plane[4 * i + 3];
                if (!(disp > 0)) continue;
                const float vdp = baseline / disp;
                const float centerX = (W - 1) * 0.5f, centerY = (W - 1) * 0.5f;
                const float offX = x - centerX, offY = y - centerY;
                const int X = (int)std::round(offX * vdp + cx);
                const int Y = (int)std::round(offY * vdp + cy);
                if (X <= cols && X >= 0 && Y <= rows && Y >= 0) {
                    const VIPoint pt{v, x, y, vdp, 0.0f};
                    co[(size_t)Y * (cols + 1) + X].push_back(pt);
                    last = pt;
                    ++ncollected;
                }
            }
    }
    // Step 2 (:285-307) with FilterMatchingPoints: score = rich/255 * 0.9 + blur/255 * 0.1 with
    // integer quotients, read at [srcImageX][srcImageY] (row = source column), std::sort by
    // descending score, first element kept.
    struct Chosen { int mi, sx, sy; float vd; bool has; };
    std::vector<Chosen> ch((size_t)N, Chosen{-1, 0, 0, 0.0f, false});
    int maxc = 0, over16 = 0;
    long long nchosen = 0;
    for (long long y = 0; y < rows; ++y)
        for (long long x = 0; x < cols; ++x) {
            std::vector<VIPoint> vec = co[(size_t)y * (cols + 1) + x];
            if (vec.empty()) continue;
            maxc = std::max(maxc, (int)vec.size());
            if (vec.size() > 16) ++over16;
            if (vec.size() > 1) {
                for (VIPoint& q : vec) {
                    const size_t t = (size_t)q.mi * WW + (size_t)q.sx * W + q.sy;
                    const float similarityScore = (float)((rich ? rich[t] : 0) / 255);
                    const float blureScore = (float)((blur ? blur[t] : 0) / 255);
                    const float a = 0.90f, b = 0.10f;
                    q.score = similarityScore * a + blureScore * b;
                }
                std::sort(vec.begin(), vec.end(),
                          [](const VIPoint& p, const VIPoint& q) { return p.score > q.score; });
            }
            const VIPoint& w = vec[0];
            ch[(size_t)(y * cols + x)] = Chosen{w.mi, w.sx, w.sy, w.vd, true};
            ++nchosen;
        }
    // Step 3: GenerateVirtualImage.
    std::vector<uint8_t> img((size_t)N * 3, 0);
    for (long long p = 0; p < N; ++p) {
        vd_out[p] = 0.0f; inv_out[p] = 0.0f; real_mm_out[p] = 0.0f;
        const Chosen& c = ch[(size_t)p];
        if (!c.has) continue;
        const float tmp = c.vd * B + bl0;
        const float real_d = (f_main * tmp) / (tmp - f_main);
        real_mm_out[p] = real_d;
        inv_out[p] = 1 / c.vd;
        vd_out[p] = c.vd;
        const uint8_t* s = bgr + ((size_t)c.mi * WW + (size_t)c.sy * W + c.sx) * 3;
        for (int k = 0; k < 3; ++k) img[(size_t)p * 3 + k] = s[k];
    }
    std::copy(img.begin(), img.end(), aif_out);
    // Step 4: HoleFilling on a copy, in place, raster order.  A filled point is recorded under
    // the MI of the last point collected in step 1 (the member m_vIPoint), with the source
    // pixel stepped away from the neighbour it came from.
    const int kstar = last.mi;
    auto hole = [&](long long p) {
        return img[p * 3] == 0 && img[p * 3 + 1] == 0 && img[p * 3 + 2] == 0;
    };
    int pixelCount = 0;
    for (long long p = 0; p < N; ++p)
        if (img[p * 3] >= 1 && img[p * 3 + 1] >= 1 && img[p * 3 + 2] >= 1) ++pixelCount;
    const int pixelCount0 = pixelCount;
    int counter = 0, passes = 0;
    const float targetFillRatio = 0.98f;
    bool bstop = true;
    while (bstop) {
        ++passes;
        int filledCounter = 0;
        for (long long y = 0; y < rows; ++y)
            for (long long x = 0; x < cols; ++x) {
                const long long p = y * cols + x;
                if (!hole(p)) continue;
                long long q;
                int nsx, nsy;
                if (x + 1 < cols && !hole(p + 1)) {
                    q = p + 1;
                    if (!ch[q].has) continue;
                    if (!(ch[q].sx - 1 >= 0)) continue;
                    nsx = ch[q].sx - 1; nsy = ch[q].sy;
                } else if (x - 1 >= 0 && !hole(p - 1)) {
                    q = p - 1;
                    if (!ch[q].has) continue;
                    if (!(ch[q].sx + 1 < W)) continue;
                    nsx = ch[q].sx + 1; nsy = ch[q].sy;
                } else if (y + 1 < rows && !hole(p + cols)) {
                    q = p + cols;
                    if (!ch[q].has) continue;
                    if (!(ch[q].sy - 1 >= 0)) continue;
                    nsx = ch[q].sx; nsy = ch[q].sy - 1;
                } else if (y - 1 >= 0 && !hole(p - cols)) {
                    q = p - cols;
                    if (!ch[q].has) continue;
                    if (!(ch[q].sy + 1 < W)) continue;
                    nsx = ch[q].sx; nsy = ch[q].sy + 1;
                } else {
                    continue;
                }
                const uint8_t* s = bgr + ((size_t)ch[q].mi * WW + (size_t)nsy * W + nsx) * 3;
                for (int k = 0; k < 3; ++k) img[p * 3 + k] = s[k];
                ++counter; ++filledCounter; ++pixelCount;
                ch[p] = Chosen{kstar, nsx, nsy, ch[p].vd, true};
            }
        const float FillRatio = static_cast<float>(pixelCount) / static_cast<float>(rows * cols);
        if (FillRatio >= targetFillRatio) bstop = false;
        else if (filledCounter == 0) bstop = false;
        else if (counter > 30000) bstop = false;
    }
    std::copy(img.begin(), img.end(), filled_out);
    // Denoise (:364-368): five 5x5 median passes of the composite (not of the filled copy).
    std::vector<uint8_t> dn(aif_out, aif_out + N * 3);
    median5_passes(dn, rows, cols, 3, 5);
    std::copy(dn.begin(), dn.end(), denoise_out);
    if (stats) {
        stats[0] = (int)ncollected; stats[1] = (int)nchosen; stats[2] = passes;
        stats[3] = counter; stats[4] = maxc; stats[5] = kstar; stats[6] = over16;
        stats[7] = pixelCount0;
    }
}

// ------------------------------------------------------------------------------------------
// Back-projection consistency fusion: LFRefocus::FuseVirtualDepth_BackProject
// (FocalStack/LFRefocus.cpp:1553-1793) with ComputeVirtualDepth (:2086-2096), ComputeVirtual3D
// (:2098-2114) and BackProject2MI (:2116-2122).  Visualisation (quantile stretch, colour maps,
// HSV boost, confidence binarisation) is display-only and not restated.

static float bp_virtual_depth(double disparity, float mla_base) {
    if (disparity <= 0) return -9999.0f;  // g_Invalid_image (Common/Common.cpp:53)
    return (float)(mla_base / disparity);
}

// points (optional): 9 floats per fused point {x, y, z, nx, ny, nz, b, g, r} in the order the
// reference pushes them; npts receives their number (points inside the frame).
static void fuse_backproject(const float* plane, const float* centers, const uint8_t* bgr,
                             const int* nbr, const int* cnt, int K, const uint8_t* garbage, int V,
                             int crop, float baseline, float B, float bl0, int raw_w, int raw_h,
                             float* vd_out, uint8_t* aif_out, uint8_t* conf_out, float* points,
                             long long* npts, long long* nfused) {
    const int W = crop, WW = crop * crop;
    const long long N = (long long)raw_w * raw_h;
    for (long long p = 0; p < N; ++p) {
        vd_out[p] = 0.0f; conf_out[p] = 0;
        aif_out[p * 3] = aif_out[p * 3 + 1] = aif_out[p * 3 + 2] = 0;
    }
    std::vector<uint8_t> masks((size_t)V * WW, 0);
    const float mla_base = baseline * bl0 / (bl0 + B);
    const float cmi = (W - 1) * 0.5f + 1.0f;  // center_mi, both axes
    auto v3d = [&](float vd, int pc, int pr, int v, float* P) {
        const float ox = pc - cmi, oy = pr - cmi;
        P[0] = ox * vd + centers[2 * v];
        P[1] = oy * vd + centers[2 * v + 1];
        P[2] = vd;
    };
    auto back = [&](const float* P, int v, float* q) {
        q[0] = ((P[0] - centers[2 * v]) / P[2]) + cmi;
        q[1] = ((P[1] - centers[2 * v + 1]) / P[2]) + cmi;
    };
    auto bad = [](float vd) { return std::fabs(vd - -9999.0f) < 1.0e-6f || !std::isfinite(vd); };
    long long np = 0, nf = 0;
    for (int a = 0; a < V; ++a) {
        if (garbage && garbage[a]) continue;
        const int n_ngb = cnt[a];
        std::vector<int> used(n_ngb);
        for (int r = 0; r < W; ++r)
            for (int c = 0; c < W; ++c) {
                if (masks[(size_t)a * WW + r * W + c] == 1) continue;
                const size_t ir = (size_t)a * WW + r * W + c;
                const float disp_ref = plane[4 * ir + 3];
                if (disp_ref < 1.0f) continue;
                const float vd_ref = bp_virtual_depth(disp_ref, mla_base);
                if (bad(vd_ref)) continue;
                float nrm[3] = {plane[4 * ir], plane[4 * ir + 1], plane[4 * ir + 2]};
                float P[3];
                v3d(vd_ref, c, r, a, P);
                const float Pref[3] = {P[0], P[1], P[2]};
                float col[3] = {(float)bgr[ir * 3], (float)bgr[ir * 3 + 1], (float)bgr[ir * 3 + 2]};
                std::fill(used.begin(), used.end(), -1);
                int num = 0;
                for (int j = 0; j < n_ngb; ++j) {
                    const int b = nbr[(size_t)a * K + j];
                    if (b < 0 || (garbage && garbage[b])) continue;
                    float q[2];
                    back(Pref, b, q);
                    const int sr = static_cast<int>(q[1] + 0.5f);
                    const int sc = static_cast<int>(q[0] + 0.5f);
                    if (sc < 0 || sc >= W || sr < 0 || sr >= W) continue;
                    if (masks[(size_t)b * WW + sr * W + sc] == 1) continue;
                    const size_t is = (size_t)b * WW + sr * W + sc;
                    const float disp_src = plane[4 * is + 3];
                    if (disp_src <= 0.0f) continue;
                    const float vd_src = bp_virtual_depth(disp_src, mla_base);
                    if (bad(vd_src)) continue;
                    float Ps[3], qr[2];
                    v3d(vd_src, sc, sr, b, Ps);
                    back(Ps, a, qr);
                    const float err = std::sqrt((c - qr[0]) * (c - qr[0]) + (r - qr[1]) * (r - qr[1]));
                    const float rel = std::fabs(vd_src - vd_ref) / std::max(vd_ref, 1.0e-6f);
                    if (err < 1.0f && rel < 0.05f) {
                        for (int k = 0; k < 3; ++k) {
                            P[k] += Ps[k];
                            nrm[k] += plane[4 * is + k];
                            col[k] += (float)bgr[is * 3 + k];
                        }
                        used[j] = sr * W + sc;
                        ++num;
                    }
                }
                if (num >= 2) {
                    // cv::Vec3f /= float: product with the float reciprocal (OpenCV matx.hpp)
                    const float denom = static_cast<float>(num) + 1.0f;
                    const float ial = 1.f / denom;
                    for (int k = 0; k < 3; ++k) { P[k] *= ial; nrm[k] *= ial; col[k] *= ial; }
                    ++nf;
                    const int vx = (int)std::nearbyint(P[0]), vy = (int)std::nearbyint(P[1]);
                    if (vy >= 0 && vy < raw_h && vx >= 0 && vx < raw_w) {
                        const long long o = (long long)vy * raw_w + vx;
                        vd_out[o] = P[2];
                        conf_out[o] = (uint8_t)std::min(255, (num + 1) * 60);
                        for (int k = 0; k < 3; ++k) {
                            const int iv = (int)std::nearbyint(col[k]);
                            aif_out[o * 3 + k] = (uint8_t)std::min(std::max(iv, 0), 255);
                        }
                        if (points) {
                            float* o9 = points + np * 9;
                            for (int k = 0; k < 3; ++k) {
                                o9[k] = P[k]; o9[3 + k] = nrm[k]; o9[6 + k] = col[k];
                            }
                        }
                        ++np;
                    }
                    for (int j = 0; j < n_ngb; ++j) {
                        if (used[j] == -1) continue;
                        masks[(size_t)nbr[(size_t)a * K + j] * WW + used[j]] = 1;
                    }
                }
            }
    }
    if (npts) *npts = np;
    if (nfused) *nfused = nf;
}

} // namespace orc

// ------------------------------------------------------------------------------------------
// C entry points (ctypes)
// ------------------------------------------------------------------------------------------
extern "C" {

void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    orc::philox(ctr, key, out);
}

int orc_layout(const orc::LayoutParams* lp, int cap, float* centers, int* origins, int* tilekeys,
               int* lens_type, int* nbr, int* cnt) {
    return orc::layout(*lp, cap, centers, origins, tilekeys, lens_type, nbr, cnt);
}

int orc_fixed_offsets(int X, int Y, int u, int v, int* out) {
    int off[12][2];
    const int n = orc::fixed_offsets(X, Y, u, v, off);
    for (int e = 0; e < n; ++e) { out[2 * e] = off[e][0]; out[2 * e + 1] = off[e][1]; }
    return n;
}

int orc_fixed_neighbors(const orc::LayoutParams* lp, int* nbr, int* cnt, int* garbage) {
    return orc::fixed_neighbors(*lp, nbr, cnt, garbage);
}

void orc_slice_u8(const uint8_t* raw, int raw_w, int raw_h, const int* origins, int V, int crop,
                  uint8_t* tiles) {
    (void)raw_h;
    for (int v = 0; v < V; ++v)
        for (int y = 0; y < crop; ++y)
            for (int x = 0; x < crop; ++x)
                tiles[((size_t)v * crop + y) * crop + x] =
                    raw[(size_t)(origins[2 * v + 1] + y) * raw_w + origins[2 * v] + x];
}

void orc_set_progress(int on) { orc::g_progress = on; }

void orc_solve_f64(const uint8_t* tiles, const double* blur, const float* centers,
                   const int* nbr, const int* cnt, int V, const orc::PmParams* p, double* plane,
                   double* cost, double* disp, uint32_t* sel, int nthreads) {
    orc::solve<double>(tiles, blur, centers, nbr, cnt, V, *p, plane, cost, disp, sel, nthreads);
}

void orc_solve_f32(const uint8_t* tiles, const double* blur, const float* centers,
                   const int* nbr, const int* cnt, int V, const orc::PmParams* p, float* plane,
                   float* cost, float* disp, uint32_t* sel, int nthreads) {
    orc::solve<float>(tiles, blur, centers, nbr, cnt, V, *p, plane, cost, disp, sel, nthreads);
}

void orc_solve_acmm_f64(const uint8_t* tiles, const double* blur, const float* centers,
                        const int* nbr, const int* cnt, int V, const orc::PmParams* p,
                        double* plane, double* cost, double* disp, uint32_t* sel, int nthreads) {
    orc::solve_acmm<double>(tiles, blur, centers, nbr, cnt, V, *p, plane, cost, disp, sel,
                            nthreads);
}

void orc_area_resize_f64(const double* src, int W, int W0, double* dst) {
    orc::area_resize<double>(src, W, W0, dst);
}

int orc_acmm_levels(int W, int* width, float* scale, int* radius, int* iterations, int* top_k) {
    orc::AcmmLevel lv[2];
    const int L = orc::acmm_levels(W, lv);
    for (int i = 0; i < L; ++i) {
        width[i] = lv[i].width; scale[i] = lv[i].scale; radius[i] = lv[i].radius;
        iterations[i] = lv[i].iterations; top_k[i] = lv[i].top_k;
    }
    return L;
}

void orc_eval_costs_f64(const uint8_t* tiles, const double* blur, const float* centers,
                        const int* nbr, const int* cnt, int V, const orc::PmParams* p,
                        const int32_t* hyp, int n, double* out, int nthreads) {
    orc::eval_costs<double>(tiles, blur, centers, nbr, cnt, V, *p, hyp, n, out, nthreads);
}

// FP64 costs with, per (hypothesis, view), the margins of NccMargins: n x K x (geo, gate,
// ferr).
void orc_eval_margins_f64(const uint8_t* tiles, const double* blur, const float* centers,
                          const int* nbr, const int* cnt, int V, const orc::PmParams* p,
                          const int32_t* hyp, int n, double* out, double* margins,
                          int nthreads) {
    orc::eval_costs<double>(tiles, blur, centers, nbr, cnt, V, *p, hyp, n, out, nthreads,
                            margins);
}

void orc_eval_costs_f32(const uint8_t* tiles, const double* blur, const float* centers,
                        const int* nbr, const int* cnt, int V, const orc::PmParams* p,
                        const int32_t* hyp, int n, float* out, int nthreads) {
    orc::eval_costs<float>(tiles, blur, centers, nbr, cnt, V, *p, hyp, n, out, nthreads);
}

void orc_filter_f32(const float* gray, const float* centers, const int* nbr, const int* cnt,
                    int V, int crop, int K, float base, const orc::FilterParams* fp,
                    const float* plane_in, const float* cost_in, const float* disp_in,
                    const uint32_t* sel_in, float* plane_out, float* cost_out, float* disp_out,
                    uint32_t* sel_out, uint8_t* reason, int nthreads) {
    orc::FilterCtx c;
    c.V = V; c.W = crop; c.K = K; c.gray = gray; c.centers = centers; c.nbr = nbr; c.cnt = cnt;
    c.base = base; c.plane = plane_in; c.cost = cost_in; c.fp = *fp;
    orc::filter_f32(c, disp_in, sel_in, plane_out, cost_out, disp_out, sel_out, reason, nthreads);
}

// The filter in the reference's own order (MIDisparityFilterCPU::FilterGlobal,
// CheckModule/MIDisparityFilter.cpp:973-985, with FilterTile's invalidation at :671-686): MIs
// one after another in map order (tile row, then column: the MI index order of the lens lattice
// here), each reading its neighbours' disparities and costs as the MIs
// before it left them and invalidating its rejected pixels before the next MI runs.  The fused
// GPU pass and orc_filter_f32 read the pre-filter snapshot instead; this order-dependent form
// measures how far the two differ.
void orc_filter_f32_inplace(const float* gray, const float* centers, const int* nbr,
                            const int* cnt, int V, int crop, int K, float base,
                            const orc::FilterParams* fp, const float* plane_in,
                            const float* cost_in, const float* disp_in, const uint32_t* sel_in,
                            float* plane_out, float* cost_out, float* disp_out, uint32_t* sel_out,
                            uint8_t* reason) {
    const size_t n = (size_t)V * crop * crop;
    std::memcpy(plane_out, plane_in, n * 4 * sizeof(float));
    std::memcpy(cost_out, cost_in, n * sizeof(float));
    orc::FilterCtx c;
    c.V = V; c.W = crop; c.K = K; c.gray = gray; c.centers = centers; c.nbr = nbr; c.cnt = cnt;
    c.base = base; c.plane = plane_out; c.cost = cost_out; c.fp = *fp;
    // one thread: MI v is filtered, and written back in place, before MI v + 1 reads it
    orc::filter_f32(c, disp_in, sel_in, plane_out, cost_out, disp_out, sel_out, reason, 1);
}

void orc_refocus_f64(const uint8_t* tiles, const float* centers, int V, float pitch, int crop,
                     int out_w, int out_h, int S, float vmin, float vmax, double* out,
                     int nthreads) {
    orc::refocus(tiles, centers, V, pitch, crop, out_w, out_h, S, vmin, vmax, out, nthreads);
}

unsigned orc_topk_f64(const double* cv, int cnt, int top_k, double* mean, int* best) {
    return orc::topk<double>(cv, cnt, top_k, *mean, *best);
}

double orc_weighted_f64(const double* cv, const double* vw, int cnt, int* argmin) {
    return orc::weighted_cost<double>(cv, vw, cnt, *argmin);
}

void orc_reencode_plane_f64(const double hn[4], int qx, int qy, double du, double dv, int px,
                            int py, double out[4]) {
    const orc::P4<double> h{hn[0], hn[1], hn[2], hn[3]};
    const orc::P4<double> o = orc::reencode_plane<double>(h, qx, qy, du, dv, px, py);
    out[0] = o.x; out[1] = o.y; out[2] = o.z; out[3] = o.w;
}

void orc_random_normals_f64(uint32_t seed, uint32_t uid, uint32_t step, int n, double* out) {
    for (int i = 0; i < n; ++i) {
        orc::Stream st(seed, uid, (uint32_t)i, step);
        orc::random_normal<double>(st, orc::SLOT_INIT_MARSAGLIA, out + 3 * (size_t)i);
    }
}

void orc_disparity_mosaic(const float* plane, const float* cost, const float* centers, int V,
                          int crop, float max_cost, int raw_w, int raw_h, float* out,
                          float* range) {
    orc::disparity_mosaic(plane, cost, centers, V, crop, max_cost, raw_w, raw_h, out, range);
}

void orc_vd_map(const float* plane, const float* centers, const float* gray, int V, int crop,
                float mla_base, int raw_w, int raw_h, float* vd, uint8_t* color, float* range) {
    orc::vd_map(plane, centers, gray, nullptr, V, crop, mla_base, raw_w, raw_h, vd, color, range);
}

// VD map with the colour crops (V x crop^2 x 3 bytes B, G, R): color is raw_h x raw_w x 3.
void orc_vd_map_bgr(const float* plane, const float* centers, const uint8_t* bgr, int V, int crop,
                    float mla_base, int raw_w, int raw_h, float* vd, uint8_t* color,
                    float* range) {
    orc::vd_map(plane, centers, nullptr, bgr, V, crop, mla_base, raw_w, raw_h, vd, color, range);
}

int orc_fit_v2r(const float* aL, const float* v, int n, double* c) {
    return orc::fit_v2r(aL, v, n, c);
}

void orc_v2r_convert(const float* v, int n, const double* c, double* z) {
    for (int i = 0; i < n; ++i) {
        const double vv = v[i];
        const double zz = (vv * c[1] + c[2]) / (1.0 - vv * c[0]);
        z[i] = (std::isnan(zz) || std::isinf(zz) || zz < 0) ? 0.0 : zz;
    }
}

// Pieces of the checkerboard update, FP64, for the pins.  Draws come from `words`, the raw
// 32-bit word of every stream slot (64 slots: views 0..14, refinement 16/17, angles 20..22,
// Marsaglia 24..55).
void orc_propagation_candidates_f64(int W, int x, int y, const double* cost, int* flag, int* pos,
                                    int* dup) {
    bool fl[8];
    orc::propagation_candidates<double>(W, x, y, cost, fl, pos, dup);
    for (int a = 0; a < 8; ++a) flag[a] = fl[a] ? 1 : 0;
}

void orc_view_priors_f64(int W, int x, int y, const uint32_t* sel, int cnt, double* prior) {
    orc::view_priors<double>(W, x, y, sel, cnt, prior);
}

// ca: 8 x 32 candidate costs (slot-major), flag: 8 ints, prior: cnt values.
void orc_view_select_f64(const double* ca, const int* flag, const double* prior, int cnt,
                         int iter, const uint32_t* words, double* prob, double* cdf,
                         double* vw, uint32_t* tsel) {
    double c[8][orc::KMAX];
    bool fl[8];
    double pr[orc::KMAX] = {0};
    for (int a = 0; a < 8; ++a) {
        fl[a] = flag[a] != 0;
        for (int b = 0; b < orc::KMAX; ++b) c[a][b] = ca[a * orc::KMAX + b];
    }
    for (int i = 0; i < cnt; ++i) pr[i] = prior[i];
    orc::Stream st(words);
    orc::view_select<double>(c, fl, pr, cnt, iter, st, vw, *tsel, prob, cdf);
}

void orc_perturbed_normal_f64(const double* h, int px, int py, double perturbation,
                              const uint32_t* words, double* out) {
    orc::Stream st(words);
    const orc::P4<double> hh{h[0], h[1], h[2], h[3]};
    const orc::P4<double> o = orc::perturbed_normal<double>(st, px, py, hh, perturbation);
    out[0] = o.x; out[1] = o.y; out[2] = o.z; out[3] = o.w;
}

void orc_refine_candidates_f64(const orc::PmParams* p, int x, int y, const double* best,
                               double bcost, const uint32_t* words, double* cand) {
    orc::Stream st(words);
    const orc::P4<double> b{best[0], best[1], best[2], best[3]};
    orc::P4<double> c[5];
    orc::refine_candidates<double>(*p, x, y, b, bcost, st, c);
    for (int i = 0; i < 5; ++i) {
        cand[4 * i] = c[i].x; cand[4 * i + 1] = c[i].y;
        cand[4 * i + 2] = c[i].z; cand[4 * i + 3] = c[i].w;
    }
}

void orc_aif_composite(const float* plane, const float* centers, const uint8_t* bgr,
                       const uint8_t* rich, const uint8_t* blur, const uint8_t* garbage, int V,
                       int crop, float baseline, float B, float bl0, float f_main, int raw_w,
                       int raw_h, float* vd, float* inv, float* real_mm, uint8_t* aif,
                       uint8_t* filled, uint8_t* denoise, int* stats) {
    orc::aif_composite(plane, centers, bgr, rich, blur, garbage, V, crop, baseline, B, bl0,
                       f_main, raw_w, raw_h, vd, inv, real_mm, aif, filled, denoise, stats);
}

void orc_fuse_backproject(const float* plane, const float* centers, const uint8_t* bgr,
                          const int* nbr, const int* cnt, int K, const uint8_t* garbage, int V,
                          int crop, float baseline, float B, float bl0, int raw_w, int raw_h,
                          float* vd, uint8_t* aif, uint8_t* conf, float* points, long long* npts,
                          long long* nfused) {
    orc::fuse_backproject(plane, centers, bgr, nbr, cnt, K, garbage, V, crop, baseline, B, bl0,
                          raw_w, raw_h, vd, aif, conf, points, npts, nfused);
}

} // extern "C"
