// lfpm_fusion.hip — the refocus-side composites of a solved frame, on the GPU.
//
// All-in-focus (AIF) composite, LFRefocus::AIFImageCompositeForMIA (FocalStack/LFRefocus.cpp:
// 183-425).  Every MI pixel with d > 0 projects to round(offset * baseline / d + centre) of a
// raw-size virtual image; each virtual pixel keeps one of its candidates by the FilterMatching-
// Points score (:2283-2301: binarised richness first, saturated blur score second, collection
// order among equals); the kept candidates give the virtual depth, its reciprocal, the real depth
// and the composite colour (GenerateVirtualImage, :2303-2328).  HoleFilling (:2362-2498) then
// fills holes in raster order from the source MIs of neighbouring pixels, and five 5x5 median
// passes of the composite give the denoised image.
//   k_aif_collect   one lane per MI pixel: projection, per-pixel atomic max of (score class,
//                   -collection index) — the first best candidate in collection order
//   k_aif_resolve   one lane per virtual pixel: depth maps, colour, hole-filling state
//   k_hole_row      one workgroup per row of a HoleFilling pass, iterated to the fixed point of
//                   the rows' dependence on the row above (the raster result); passes until the
//                   reference's stop rule holds
//
// Back-projection consistency fusion, LFRefocus::FuseVirtualDepth_BackProject (:1553-1793).
// Every MI pixel's virtual point is back-projected into its matching neighbours; a neighbour
// pixel whose own virtual point re-projects within 1 px and 5 % virtual depth is consistent; with
// at least two consistent neighbours not yet used, the point, normal and colour are averaged and
// splatted, and the neighbour pixels used are masked for the rest of the walk.  The walk is
// sequential in the reference (MI order, pixel order); here
//   k_bp_consistency   one lane per MI pixel: the consistency bits of all neighbours (masks
//                      only remove candidates, so they are applied afterwards)
//   k_bp_level_wg      one workgroup per MI, level by level of a schedule in which MIs of one
//                      level share no mask (lfpm_fuse_schedule): the masked walk over the MI's
//                      pixels as the fixed point of its triangular system, thread = pixel
//   k_bp_level         the same walk step by step, one wavefront per MI, lane = neighbour
//                      (crops above 32)
//   k_bp_emit          one lane per fused pixel: the averages in neighbour order, splat by the
//                      highest walk index (the reference's last write), point records in order
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "../../include/lfpm.h"
#include "lfpm_internal.h"

namespace lfpm {
namespace {

constexpr uint32_t kSeqMask = 0x3FFFFFFFu;   // collection index field of a candidate key
// cc word flags of the hole-filling row state (colour in bits 0..23)
constexpr uint32_t kDep = 1u << 31;      // outcome depends on the left neighbour
constexpr uint32_t kDFill = 1u << 30;    // the down / up alternative fills
constexpr uint32_t kChanged = 1u << 29;  // the pixel was filled
constexpr uint32_t kRgb = 0x00FFFFFFu;
constexpr float kInvalidVd = -9999.0f;   // g_Invalid_image (Common/Common.cpp:53)

size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

int grid_for(long long n, int threads) {
    long long g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > 0x7fffffffLL) g = 0x7fffffffLL;
    return (int)g;
}

int check_launch(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(LFPM_E_HIP, "%s: %s", what, hipGetErrorString(e));
    return LFPM_OK;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, long long n, uint32_t v) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_u8(uint8_t* __restrict__ p, long long n) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Map word of a virtual pixel's source: (MI + 1) << 14 | column << 7 | row; 0 = none.
__device__ __forceinline__ uint32_t map_word(int mi, int sx, int sy) {
    return ((uint32_t)(mi + 1) << 14) | ((uint32_t)sx << 7) | (uint32_t)sy;
}
__device__ __forceinline__ int map_mi(uint32_t m) { return (int)(m >> 14) - 1; }
__device__ __forceinline__ int map_sx(uint32_t m) { return (int)((m >> 7) & 127u); }
__device__ __forceinline__ int map_sy(uint32_t m) { return (int)(m & 127u); }

__device__ __forceinline__ uint32_t rgb_of(const uint8_t* __restrict__ bgr, long long px) {
    const uint8_t* s = bgr + px * 3;
    return (uint32_t)s[0] | ((uint32_t)s[1] << 8) | ((uint32_t)s[2] << 16);
}

// ---------------------------------------------------------------------------------------------
// AIF composite

// stats: [0] collected points, [1] chosen pixels, [2] hole-filling passes, [3] fills, [4] most
// candidates at one pixel, [5] the MI filled points are attributed to, [6] pixels with more than
// 16 candidates, [7] pixels with every channel >= 1 before filling.
__global__ void __launch_bounds__(256) k_aif_collect(
    const float4* __restrict__ plane, const float2* __restrict__ centers,
    const uint8_t* __restrict__ rich, const uint8_t* __restrict__ blur,
    const uint8_t* __restrict__ garbage, int V, int W, float baseline, int cols, int rows,
    uint32_t* __restrict__ win, int* __restrict__ cand, int* __restrict__ stats) {
#pragma clang fp contract(off)
    const int WW = W * W;
    const long long M = (long long)V * WW;
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    bool collected = false;
    if (i < M) {
        const int v = (int)(i / WW), pix = (int)(i - (long long)v * WW);
        const int x = pix % W, y = pix / W;
        const float disp = plane[i].w;
        if (!(garbage && garbage[v]) && disp > 0.0f) {
            const float vdp = baseline / disp;
            const float centre = (float)(W - 1) * 0.5f;
            const float offX = (float)x - centre, offY = (float)y - centre;
            const float2 c = centers[v];
            const float fx = offX * vdp + c.x, fy = offY * vdp + c.y;
            const int X = (int)roundf(fx), Y = (int)roundf(fy);
            if (X >= 0 && X <= cols && Y >= 0 && Y <= rows) {
                collected = true;
                if (X < cols && Y < rows) {
                    // FilterMatchingPoints reads the score maps at [srcImageX][srcImageY]:
                    // row = source column
                    const long long t = (long long)v * WW + (long long)x * W + y;
                    const int s = (rich ? rich[t] : 0) / 255, b = (blur ? blur[t] : 0) / 255;
                    const uint32_t key = ((uint32_t)(2 * s + b) << 30) | (kSeqMask - (uint32_t)i);
                    const long long p = (long long)Y * cols + X;
                    atomicMax(&win[p], key);
                    atomicAdd(&cand[p], 1);
                }
            }
        }
    }
    // block totals: candidates collected, and the last one collected (highest index) for the
    // attribution MI
    __shared__ int s_n, s_last;
    if (threadIdx.x == 0) { s_n = 0; s_last = -1; }
    __syncthreads();
    const unsigned long long m = __ballot(collected);
    const int lane = threadIdx.x & 63;
    if (m) {
        if (lane == __ffsll((long long)m) - 1) atomicAdd(&s_n, __popcll(m));
        if (lane == 63 - __clzll((long long)m)) atomicMax(&s_last, (int)i);
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_n) {
        atomicAdd(&stats[0], s_n);
        atomicMax(&stats[5], s_last);
    }
}

__global__ void __launch_bounds__(256) k_aif_resolve(
    const float4* __restrict__ plane, const uint8_t* __restrict__ bgr, const uint32_t* __restrict__ win,
    const int* __restrict__ cand, int W, float baseline, float B, float bl0, float f_main,
    long long N, float* __restrict__ vd_out, float* __restrict__ inv_out,
    float* __restrict__ real_out, uint8_t* __restrict__ aif, uint32_t* __restrict__ col32,
    uint32_t* __restrict__ map32, int* __restrict__ stats) {
#pragma clang fp contract(off)
    const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    bool chosen = false, over16 = false;
    int nc = 0;
    if (p < N) {
        const uint32_t w = win[p];
        nc = cand[p];
        over16 = nc > 16;
        if (w == 0) {
            vd_out[p] = 0.0f; inv_out[p] = 0.0f; real_out[p] = 0.0f;
            aif[p * 3] = aif[p * 3 + 1] = aif[p * 3 + 2] = 0;
            col32[p] = 0; map32[p] = 0;
        } else {
            chosen = true;
            const long long seq = (long long)(kSeqMask - (w & kSeqMask));
            const int WW = W * W;
            const int v = (int)(seq / WW), pix = (int)(seq - (long long)v * WW);
            const float vdp = baseline / plane[seq].w;
            const float tmp = vdp * B + bl0;
            const float real_d = (f_main * tmp) / (tmp - f_main);
            vd_out[p] = vdp;
            inv_out[p] = 1.0f / vdp;
            real_out[p] = real_d;
            const uint32_t c = rgb_of(bgr, seq);
            aif[p * 3] = (uint8_t)c; aif[p * 3 + 1] = (uint8_t)(c >> 8);
            aif[p * 3 + 2] = (uint8_t)(c >> 16);
            col32[p] = c;
            map32[p] = map_word(v, pix % W, pix / W);
        }
    }
    __shared__ int s_c, s_o, s_m;
    if (threadIdx.x == 0) { s_c = 0; s_o = 0; s_m = 0; }
    __syncthreads();
    const unsigned long long mc = __ballot(chosen), mo = __ballot(over16);
    const int lane = threadIdx.x & 63;
    // wave maximum of the candidate counts
    int mx = nc;
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o));
    if (lane == 0) {
        if (mc) atomicAdd(&s_c, __popcll(mc));
        if (mo) atomicAdd(&s_o, __popcll(mo));
        if (mx > 0) atomicMax(&s_m, mx);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_c) atomicAdd(&stats[1], s_c);
        if (s_o) atomicAdd(&stats[6], s_o);
        if (s_m) atomicMax(&stats[4], s_m);
    }
}

// The fill HoleFilling takes for pixel p from neighbour word m (its source, `dsx`/`dsy` the step
// to the filled source pixel); false when the reference skips the pixel.
__device__ __forceinline__ bool fill_from(uint32_t m, int dsx, int dsy, int W,
                                          const uint8_t* __restrict__ bgr,
                                          const uint32_t* kcrop, int kstar, uint32_t& col,
                                          uint32_t& map) {
    if (m == 0) return false;  // a pixel not in the point map: `continue`
    const int sx = map_sx(m) + dsx, sy = map_sy(m) + dsy;
    if (sx < 0 || sx >= W || sy < 0 || sy >= W) return false;
    const int mi = map_mi(m);
    col = (kcrop && mi == kstar) ? kcrop[sy * W + sx]
                                 : rgb_of(bgr, (long long)mi * W * W + (long long)sy * W + sx);
    map = map_word(kstar, sx, sy);
    return true;
}

// HoleFilling (FocalStack/LFRefocus.cpp:2362-2498).  Raster order within a pass: a pixel reads
// its right and lower neighbours as they were when the pass started (A) and its left and upper
// neighbours as already updated.  Row y therefore depends on the pass-start state and on the
// updated row y - 1, and on the latter only through the "up" option (a hole whose right, left-at-
// visit and lower neighbours are holes).  A pass runs as a fixed-point iteration over rows:
// iteration 1 computes every row against A for the row above; iteration t recomputes each row
// that consulted the row above and whose row above changed since, against that row's latest
// result.  Row 0 is final after iteration 1 and row y after iteration y + 1 at the latest; the
// fixed point is the sequential raster result.  Each row keeps two result slots (the slot its
// lower neighbour reads in an iteration is never the one written), a version counting its
// changes, and the version of the row above it was computed against.  Within a row, pixels whose
// outcome does not depend on the left neighbour are decided in parallel; the runs that do (holes
// with a hole to the right) are walked by one lane each.
struct HoleState {
    const uint32_t* a_col;  // pass-start colours (B, G, R packed; 0 = hole)
    const uint32_t* a_map;  // pass-start source words
    uint32_t* b_col0; uint32_t* b_map0; uint32_t* b_col1; uint32_t* b_map1;  // result slots
    int* slot;   // [2][rows] by iteration parity: 0 / 1 = result slot, 2 = the pass-start row
    int* ver;    // [2][rows] versions
    int* dep;    // [rows] version of the row above used; -1 = the row above was not consulted
    int* fills;  // [rows] fills of the row's latest result
    int* dirty;  // [iterations + 1] rows recomputed per iteration
};

constexpr int kRowThreads = 256;

__global__ void __launch_bounds__(kRowThreads) k_hole_row(HoleState h,
                                                          const uint8_t* __restrict__ bgr,
                                                          int W, int cols, int rows,
                                                          const int* __restrict__ stats,
                                                          int stage_kcrop, int t) {
    extern __shared__ __align__(16) uint32_t lds[];
    uint32_t* cc = lds;             // row state: colour | flags
    uint32_t* cm = lds + cols;      // row state: source word
    uint32_t* kl = lds + 2 * cols;  // colours of the attribution MI's crop (when staged)
    uint8_t* kz = (uint8_t*)(kl + W * W);  // [sy][s]: first column >= s of row sy that is black
    __shared__ int s_go, s_changed, s_fills, s_consult;
    const int y = blockIdx.x, tid = threadIdx.x;
    const int rd = (t + 1) & 1, wr = t & 1;
    const int sp = h.slot[rd * rows + y], vp = h.ver[rd * rows + y];
    if (tid == 0) {
        int go = t == 1;
        if (!go && y > 0 && h.dep[y] >= 0 && h.ver[rd * rows + y - 1] != h.dep[y]) go = 1;
        s_go = go;
        s_changed = 0;
        s_fills = 0;
        s_consult = 0;
        if (!go) {
            h.slot[wr * rows + y] = sp;
            h.ver[wr * rows + y] = vp;
        }
    }
    __syncthreads();
    if (!s_go) return;
    const int kstar = stats[5];
    const uint32_t* kcrop = nullptr;
    if (stage_kcrop && kstar >= 0) {
        for (int i = tid; i < W * W; i += kRowThreads)
            kl[i] = rgb_of(bgr, (long long)kstar * W * W + i);
        kcrop = kl;
    }
    __syncthreads();
    if (kcrop) {
        for (int r = tid; r < W; r += kRowThreads) {
            int z = W;
            for (int c = W - 1; c >= 0; --c) {
                if (kl[r * W + c] == 0) z = c;
                kz[r * W + c] = (uint8_t)z;
            }
        }
    }
    const int su = y > 0 ? h.slot[rd * rows + y - 1] : 2;
    const int vu = y > 0 ? h.ver[rd * rows + y - 1] : 0;
    const uint32_t* up_col = su == 2 ? h.a_col : (su == 0 ? h.b_col0 : h.b_col1);
    const uint32_t* up_map = su == 2 ? h.a_map : (su == 0 ? h.b_map0 : h.b_map1);
    const long long row = (long long)y * cols;
    __syncthreads();
    bool consult = false;
    for (int x = tid; x < cols; x += kRowThreads) {
        const long long p = row + x;
        const uint32_t c0 = h.a_col[p];
        uint32_t oc, om = 0;
        if (c0 != 0) {
            oc = c0;
            om = h.a_map[p];
        } else if (x + 1 < cols && h.a_col[p + 1] != 0) {
            uint32_t c, m;
            if (fill_from(h.a_map[p + 1], -1, 0, W, bgr, kcrop, kstar, c, m)) {
                oc = c | kChanged;
                om = m;
            } else {
                oc = 0;
            }
        } else {
            uint32_t c = 0, m = 0;
            bool f = false;
            if (y + 1 < rows && h.a_col[p + cols] != 0) {
                f = fill_from(h.a_map[p + cols], 0, -1, W, bgr, kcrop, kstar, c, m);
            } else if (y >= 1) {
                consult = true;
                if (up_col[p - cols] != 0)
                    f = fill_from(up_map[p - cols], 0, 1, W, bgr, kcrop, kstar, c, m);
            }
            oc = kDep | (f ? (kDFill | c) : 0u);
            om = f ? m : 0u;
        }
        cc[x] = oc;
        cm[x] = om;
    }
    if (consult) s_consult = 1;
    __syncthreads();
    int fills = 0;
    // runs of left-dependent pixels, one lane each.  A pixel with a non-hole left neighbour
    // fills from the left (the source one column right of the neighbour's); a chain of such
    // fills inside the attribution MI's crop runs along one crop row until the crop edge or a
    // black source pixel, so it is written in one stretch from the first-black table.
    for (int x = tid; x < cols; x += kRowThreads) {
        if (!(cc[x] & kDep) || (x > 0 && (cc[x - 1] & kDep))) continue;
        int x1 = x + 1;
        while (x1 < cols && (cc[x1] & kDep)) ++x1;
        bool lnon = x > 0 && (cc[x - 1] & kRgb) != 0;
        uint32_t lm = x > 0 ? cm[x - 1] : 0u;
        int xx = x;
        while (xx < x1) {
            if (lnon && kcrop && lm != 0 && map_mi(lm) == kstar) {
                const int sx = map_sx(lm), sy = map_sy(lm);
                if (sx + 1 >= W) {  // no source to the right: the pixel stays a hole
                    cc[xx] = kDep;
                    lnon = false;
                    ++xx;
                    continue;
                }
                const int z = kz[sy * W + sx + 1];  // first black column >= sx + 1 (W: none)
                const int last = min(z, W - 1);     // last source column of the stretch
                const int n = min(last - sx, x1 - xx);
                for (int k = 0; k < n; ++k) {
                    const int col = sx + 1 + k;
                    cc[xx + k] = kDep | kChanged | kl[sy * W + col];
                    cm[xx + k] = map_word(kstar, col, sy);
                }
                xx += n;
                fills += n;
                lm = map_word(kstar, sx + n, sy);
                lnon = kl[sy * W + sx + n] != 0;
                continue;
            }
            const uint32_t d = cc[xx];
            uint32_t c = 0, m = 0;
            bool f;
            if (lnon) {
                f = fill_from(lm, 1, 0, W, bgr, kcrop, kstar, c, m);
            } else if (d & kDFill) {
                f = true;
                c = d & kRgb;
                m = cm[xx];
            } else {
                f = false;
            }
            if (f) {
                cc[xx] = kDep | kChanged | c;
                cm[xx] = m;
                ++fills;
            } else {
                cc[xx] = kDep;
            }
            lnon = f && c != 0;
            lm = m;
            ++xx;
        }
    }
    __syncthreads();
    const int ns = sp == 0 ? 1 : 0;
    uint32_t* out_col = ns == 0 ? h.b_col0 : h.b_col1;
    uint32_t* out_map = ns == 0 ? h.b_map0 : h.b_map1;
    const uint32_t* prev_col = sp == 2 ? h.a_col : (sp == 0 ? h.b_col0 : h.b_col1);
    const uint32_t* prev_map = sp == 2 ? h.a_map : (sp == 0 ? h.b_map0 : h.b_map1);
    bool changed = false;
    for (int x = tid; x < cols; x += kRowThreads) {
        const long long p = row + x;
        const uint32_t d = cc[x];
        const bool f = (d & kChanged) != 0;
        const uint32_t nc = f ? (d & kRgb) : h.a_col[p];
        const uint32_t nm = f ? cm[x] : h.a_map[p];
        out_col[p] = nc;
        out_map[p] = nm;
        fills += f && !(d & kDep);
        changed |= nc != prev_col[p] || nm != prev_map[p];
    }
    if (changed) s_changed = 1;
    atomicAdd(&s_fills, fills);
    __syncthreads();
    if (tid == 0) {
        h.slot[wr * rows + y] = ns;
        h.ver[wr * rows + y] = vp + (s_changed ? 1 : 0);
        h.dep[y] = s_consult ? vu : -1;
        h.fills[y] = s_fills;
        atomicAdd(&h.dirty[t], 1);
    }
}

// End of a pass: the rows' latest results become the pass-start state of the next pass.
__global__ void __launch_bounds__(256) k_hole_gather(HoleState h, int cols, int rows, int par,
                                                     uint32_t* __restrict__ col32,
                                                     uint32_t* __restrict__ map32) {
    const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (p >= (long long)cols * rows) return;
    const int y = (int)(p / cols);
    const int sl = h.slot[par * rows + y];
    if (sl == 2) return;
    col32[p] = sl == 0 ? h.b_col0[p] : h.b_col1[p];
    map32[p] = sl == 0 ? h.b_map0[p] : h.b_map1[p];
}

// pass state: [0] pixel count, [1] fills in all passes, [2] passes, [3] stop
__global__ void __launch_bounds__(256) k_hole_count(const uint32_t* __restrict__ col32, long long N,
                                                    int* __restrict__ pstate) {
    const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    bool in = false;
    if (p < N) {
        const uint32_t c = col32[p];
        in = (c & 0xFFu) >= 1 && ((c >> 8) & 0xFFu) >= 1 && ((c >> 16) & 0xFFu) >= 1;
    }
    __shared__ int s_n;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const unsigned long long m = __ballot(in);
    if (m && (threadIdx.x & 63) == __ffsll((long long)m) - 1) atomicAdd(&s_n, __popcll(m));
    __syncthreads();
    if (threadIdx.x == 0 && s_n) atomicAdd(&pstate[0], s_n);
}

__global__ void k_hole_pass_end(const int* __restrict__ fills, int rows, int cols,
                                int* __restrict__ pstate, int* __restrict__ stats) {
    __shared__ int s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    int f = 0;
    for (int y = threadIdx.x; y < rows; y += blockDim.x) f += fills[y];
    atomicAdd(&s, f);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int fl = s;
        pstate[0] += fl;
        pstate[1] += fl;
        pstate[2] += 1;
        const float ratio = static_cast<float>(pstate[0]) / static_cast<float>(rows * cols);
        pstate[3] = (ratio >= 0.98f) || fl == 0 || pstate[1] > 30000;
        stats[2] = pstate[2];
        stats[3] = pstate[1];
    }
}

__global__ void k_hole_init(int* __restrict__ slot0, int* __restrict__ ver0, int rows,
                            int* __restrict__ dirty, int niter) {
    for (int y = threadIdx.x; y < rows; y += blockDim.x) {
        slot0[y] = 2;
        ver0[y] = 0;
    }
    for (int i = threadIdx.x; i < niter; i += blockDim.x) dirty[i] = 0;
}

// The attribution MI (the collection index of the last candidate collected -> its MI) and the
// pixel count before filling.
__global__ void k_aif_kstar(int* __restrict__ stats, int WW, int* __restrict__ pstate) {
    if (stats[5] >= 0) stats[5] /= WW;
    stats[7] = pstate[0];
}

__global__ void __launch_bounds__(256) k_unpack_rgb(const uint32_t* __restrict__ col32, long long N,
                                                    uint8_t* __restrict__ out) {
    const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (p >= N) return;
    const uint32_t c = col32[p];
    out[p * 3] = (uint8_t)c; out[p * 3 + 1] = (uint8_t)(c >> 8); out[p * 3 + 2] = (uint8_t)(c >> 16);
}

__global__ void __launch_bounds__(256) k_split3(const uint8_t* __restrict__ in, long long N,
                                                uint8_t* __restrict__ planes) {
    const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (p >= N) return;
    planes[p] = in[p * 3]; planes[N + p] = in[p * 3 + 1]; planes[2 * N + p] = in[p * 3 + 2];
}

struct AifLayout {
    size_t win, cand, col32, map32, planes, tmp0, tmp1, b[4], slot, ver, dep, fills, dirty,
        pstate, total;
};

constexpr int kHoleGroup = 8;  // iterations launched between two reads of their dirty counts

AifLayout aif_layout(int raw_w, int raw_h) {
    const size_t N = (size_t)raw_w * raw_h, R = (size_t)raw_h;
    AifLayout L;
    size_t o = 0;
    L.win = o; o += al(N * 4);
    L.cand = o; o += al(N * 4);
    L.col32 = o; o += al(N * 4);
    L.map32 = o; o += al(N * 4);
    L.planes = o; o += al(N * 3);
    L.tmp0 = o; o += al(N);
    L.tmp1 = o; o += al(N);
    for (int i = 0; i < 4; ++i) { L.b[i] = o; o += al(N * 4); }
    L.slot = o; o += al(2 * R * 4);
    L.ver = o; o += al(2 * R * 4);
    L.dep = o; o += al(R * 4);
    L.fills = o; o += al(R * 4);
    L.dirty = o; o += al((R + 2 + kHoleGroup) * 4);
    L.pstate = o; o += al(4 * 4);
    L.total = o;
    return L;
}

// ---------------------------------------------------------------------------------------------
// Back-projection consistency fusion

struct BpGeom {
    float mla_base;  // baseline * bl0 / (bl0 + B): the microlens baseline
    float cmi;       // (crop - 1) * 0.5 + 1: the centre ComputeVirtual3D / BackProject2MI use
    int W;
};

// ComputeVirtualDepth (FocalStack/LFRefocus.cpp:2086-2096): the disparity enters as a double.
__device__ __forceinline__ float bp_vd(float disp, float mla_base) {
    if (disp <= 0.0f) return kInvalidVd;
    return (float)((double)mla_base / (double)disp);
}
__device__ __forceinline__ bool bp_bad(float vd) {
    return fabsf(vd - kInvalidVd) < 1.0e-6f || !isfinite(vd);
}

// Source pixel of the reference point (vd, c, r) of an MI centred at ca in the neighbour
// centred at cb (BackProject2MI, rounded by +0.5 and truncation); false when outside the crop.
__device__ __forceinline__ bool bp_source(float vd, int c, int r, float2 ca, float2 cb,
                                          const BpGeom& g, int& sr, int& sc) {
#pragma clang fp contract(off)
    const float P0 = ((float)c - g.cmi) * vd + ca.x;
    const float P1 = ((float)r - g.cmi) * vd + ca.y;
    const float q0 = ((P0 - cb.x) / vd) + g.cmi;
    const float q1 = ((P1 - cb.y) / vd) + g.cmi;
    sr = static_cast<int>(q1 + 0.5f);
    sc = static_cast<int>(q0 + 0.5f);
    return !(sc < 0 || sc >= g.W || sr < 0 || sr >= g.W);
}

// cons[a*WW + p]: bit 31 = the reference pixel enters the walk (d >= 1, valid virtual depth),
// bit j = neighbour j is consistent (ignoring masks).
__global__ void __launch_bounds__(256) k_bp_consistency(
    const float4* __restrict__ plane, const float2* __restrict__ centers,
    const int* __restrict__ nbr, const int* __restrict__ cnt, int K,
    const uint8_t* __restrict__ garbage, int V, BpGeom g, uint32_t* __restrict__ cons,
    unsigned short* __restrict__ srcs) {
#pragma clang fp contract(off)
    const int W = g.W, WW = W * W;
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= (long long)V * WW) return;
    const int a = (int)(i / WW), pix = (int)(i - (long long)a * WW);
    const int c = pix % W, r = pix / W;
    uint32_t bits = 0;
    const float disp_ref = plane[i].w;
    if (!(garbage && garbage[a]) && disp_ref >= 1.0f) {
        const float vd_ref = bp_vd(disp_ref, g.mla_base);
        if (!bp_bad(vd_ref)) {
            bits = 1u << 31;
            const float2 ca = centers[a];
            const int n = min(cnt[a], K);
            for (int j = 0; j < n; ++j) {
                const int b = nbr[(long long)a * K + j];
                if (b < 0 || (garbage && garbage[b])) continue;
                const float2 cb = centers[b];
                int sr, sc;
                if (!bp_source(vd_ref, c, r, ca, cb, g, sr, sc)) continue;
                const float disp_src = plane[(long long)b * WW + sr * W + sc].w;
                if (disp_src <= 0.0f) continue;
                const float vd_src = bp_vd(disp_src, g.mla_base);
                if (bp_bad(vd_src)) continue;
                const float Ps0 = ((float)sc - g.cmi) * vd_src + cb.x;
                const float Ps1 = ((float)sr - g.cmi) * vd_src + cb.y;
                const float r0 = ((Ps0 - ca.x) / vd_src) + g.cmi;
                const float r1 = ((Ps1 - ca.y) / vd_src) + g.cmi;
                const float e0 = (float)c - r0, e1 = (float)r - r1;
                const float err = sqrtf(e0 * e0 + e1 * e1);
                const float rel = fabsf(vd_src - vd_ref) / fmaxf(vd_ref, 1.0e-6f);
                if (err < 1.0f && rel < 0.05f) {
                    bits |= 1u << j;
                    srcs[i * K + j] = (unsigned short)(sr * W + sc);
                }
            }
        }
    }
    cons[i] = bits;
}

// One wavefront per MI of one schedule level: the reference's masked walk over the MI's pixels.
// masks: per MI ceil(WW/32) words.  incl[a*WW + p] = the neighbours fused at pixel p (0: none).
// Masks only remove candidates, so pixels with fewer than two consistent neighbours never fuse:
// the walk visits only the others (in pixel order), with their source pixels computed in
// parallel beforehand.
__global__ void __launch_bounds__(64) k_bp_level(
    const int* __restrict__ order, int begin, const float4* __restrict__ plane,
    const float2* __restrict__ centers, const int* __restrict__ nbr, const int* __restrict__ cnt,
    int K, const uint8_t* __restrict__ garbage, BpGeom g, const uint32_t* __restrict__ cons,
    const unsigned short* __restrict__ srcs, uint32_t* __restrict__ masks,
    uint32_t* __restrict__ incl) {
    extern __shared__ __align__(16) uint32_t lds[];
    const int W = g.W, WW = W * W, MW = (WW + 31) >> 5;
    uint32_t* ml = lds;                                  // masks: own, then entry j at (1 + j)*MW
    int* pl = (int*)(lds + (1 + K) * MW);                // walk list: pixel indices
    uint32_t* cl = (uint32_t*)(pl + WW);                 // consistency bits by pixel
    unsigned short* sl = (unsigned short*)(cl + WW);     // source pixel of pixel p, neighbour j
    __shared__ int s_nb[32], s_slot[32];
    const int a = order[begin + blockIdx.x];
    const int lane = threadIdx.x;
    const int n = max(0, min(cnt[a], K));
    int b = -1;
    if (lane < n) {
        b = nbr[(long long)a * K + lane];
        if (b >= 0 && garbage && garbage[b]) b = -1;
    }
    // a neighbour listed twice shares the mask copy of its first entry
    int slot = lane;
    for (int j = 0; j < n; ++j) {
        const int bj = __shfl(b, j);
        if (j < slot && bj == b && b >= 0) slot = j;
    }
    if (lane < n) {
        s_nb[lane] = b;
        s_slot[lane] = slot;
    }
    const long long base = (long long)a * WW;
#pragma unroll 4
    for (int t = lane; t < WW; t += 64) cl[t] = cons[base + t];
    __syncthreads();
    for (int q = lane; q < (1 + n) * MW; q += 64) {
        const int e = q / MW, t = q - e * MW;
        const int mi = e == 0 ? a : s_nb[e - 1];
        if (mi < 0 || (e > 0 && s_slot[e - 1] != e - 1)) continue;
        ml[q] = masks[(long long)mi * MW + t];
    }
    // walk list: pixels entering the walk with at least two consistent neighbours, in order
    int len = 0;
    const uint32_t nmask = (1u << n) - 1u;
    for (int c0 = 0; c0 < WW; c0 += 64) {
        const int p = c0 + lane;
        const uint32_t cb = p < WW ? cl[p] : 0u;
        const bool take = (cb >> 31) && __popc(cb & nmask) >= 2;
        const unsigned long long m = __ballot(take);
        if (take) pl[len + __popcll(m & ((1ull << lane) - 1ull))] = p;
        len += __popcll(m);
    }
    // source pixels of every (pixel, consistent neighbour) from the consistency pass: the MI's
    // block of WW * K entries, copied with 16 loads in flight per lane
    {
        const int total = WW * K;
        const unsigned short* src = srcs + base * K;
        for (int q0 = lane; q0 < total; q0 += 64 * 16) {
            unsigned short v[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int q = q0 + 64 * u;
                v[u] = q < total ? src[q] : (unsigned short)0;
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int q = q0 + 64 * u;
                if (q < total) sl[q] = v[u];
            }
        }
    }
    __syncthreads();
    bool dup = false;
    for (int j = 0; j < n; ++j) dup |= s_slot[j] != j && s_nb[j] >= 0;
    if (MW <= 8 && !dup) {
        // masks in registers: every lane the MI's own, lane j neighbour j's (the own mask does
        // not change during the MI's walk); the next entry is read while this one is decided
        uint32_t own[8], nm[8];
#pragma unroll
        for (int w = 0; w < 8; ++w) {
            own[w] = w < MW ? ml[w] : 0u;
            nm[w] = (w < MW && lane < n && b >= 0) ? ml[(1 + lane) * MW + w] : 0u;
        }
        int pn = len > 0 ? pl[0] : 0;
        uint32_t cn = len > 0 ? cl[pn] : 0u;
        int sn = (len > 0 && lane < n) ? sl[pn * K + lane] : 0;
        for (int i = 0; i < len; ++i) {
            const int p = pn;
            const uint32_t cbits = cn;
            const int sidx = sn;
            if (i + 1 < len) {
                pn = pl[i + 1];
                cn = cl[pn];
                sn = lane < n ? sl[pn * K + lane] : 0;
            }
            uint32_t ow = 0;
#pragma unroll
            for (int w = 0; w < 8; ++w) ow = (p >> 5) == w ? own[w] : ow;
            if ((ow >> (p & 31)) & 1u) continue;
            bool inc = false;
            if (lane < n && ((cbits >> lane) & 1u)) {
                uint32_t mw = 0;
#pragma unroll
                for (int w = 0; w < 8; ++w) mw = (sidx >> 5) == w ? nm[w] : mw;
                inc = !((mw >> (sidx & 31)) & 1u);
            }
            const unsigned long long m = __ballot(inc);
            if (__popcll(m) >= 2) {
                if (lane == 0) incl[base + p] = (uint32_t)m;
                if (inc) {
#pragma unroll
                    for (int w = 0; w < 8; ++w)
                        if ((sidx >> 5) == w) nm[w] |= 1u << (sidx & 31);
                }
            }
        }
        if (lane < n && b >= 0) {
#pragma unroll
            for (int w = 0; w < 8; ++w)
                if (w < MW) ml[(1 + lane) * MW + w] = nm[w];
        }
    } else {
        for (int i = 0; i < len; ++i) {
            const int p = pl[i];
            if ((ml[p >> 5] >> (p & 31)) & 1u) continue;
            const uint32_t cbits = cl[p];
            bool inc = false;
            int sidx = 0;
            if (lane < n && ((cbits >> lane) & 1u)) {
                sidx = sl[p * K + lane];
                inc = !((ml[(1 + slot) * MW + (sidx >> 5)] >> (sidx & 31)) & 1u);
            }
            const unsigned long long m = __ballot(inc);
            if (__popcll(m) >= 2) {
                if (lane == 0) incl[base + p] = (uint32_t)m;
                if (inc) atomicOr(&ml[(1 + slot) * MW + (sidx >> 5)], 1u << (sidx & 31));
                __syncthreads();
            }
        }
    }
    __syncthreads();
    for (int q = lane + MW; q < (1 + n) * MW; q += 64) {
        const int e = q / MW, t = q - e * MW;
        const int mi = s_nb[e - 1];
        if (mi < 0 || s_slot[e - 1] != e - 1) continue;
        masks[(long long)mi * MW + t] = ml[q];
    }
}

constexpr int kBpMaxPix = 1024;  // crop^2 bound of the workgroup walk: one thread per pixel
constexpr int kBpMaxNb = 31;     // neighbour bits of a consistency word (bit 31: in the walk)

// One workgroup per MI of one schedule level, one thread per MI pixel: the masked walk decided
// for all of the MI's pixels at once.  A pixel's outcome depends on earlier pixels only through
// the neighbour pixels their fusions mark, so the walk is the unique solution of the triangular
// system
//   inc(p)   = inc0(p) minus {j : a fused q < p marks the source pixel of p in neighbour j},
//   fused(p) = |inc(p)| >= 2,
// where inc0(p) holds the consistent neighbours whose source pixel is unmarked when the MI's walk
// starts.  Rounds of Jacobi iteration from fused = |inc0| >= 2 reach it (after k rounds the first
// k walk pixels are final) and a round that changes nothing ends them, since a fixed point is
// that solution.  Only (mask copy, source pixel) pairs that two walk entries target can change
// an outcome, so the rounds visit only those.
// A pair is the bit index slot * MW * 32 + source of the neighbour masks after the MI's own
// (each copy padded to whole words).  claim[pair]: the number of entries targeting it, then per
// round r the key r << 16 | (0xFFFF - q) of the first fused q that targets it (atomic max: the
// newest round, then the smallest q).
__global__ void __launch_bounds__(kBpMaxPix) k_bp_level_wg(
    const int* __restrict__ order, int begin, const int* __restrict__ nbr,
    const int* __restrict__ cnt, int K, const uint8_t* __restrict__ garbage, int W,
    const uint32_t* __restrict__ cons, const unsigned short* __restrict__ srcs,
    uint32_t* __restrict__ masks, uint32_t* __restrict__ incl) {
    extern __shared__ __align__(16) uint32_t lds[];
    const int WW = W * W, MW = (WW + 31) >> 5;
    uint32_t* ml = lds;                    // masks: own, then entry j at (1 + j) * MW
    uint32_t* nm = lds + MW;               // the neighbour copies as one bit array
    uint32_t* claim = lds + (1 + K) * MW;  // K * MW * 32
    __shared__ int s_nb[32], s_slot[32];
    __shared__ int s_any, s_chg[2];
    const int a = order[begin + blockIdx.x];
    const int p = threadIdx.x;
    const int n = max(0, min(cnt[a], K));
    if (p < 64) {
        int b = -1;
        if (p < n) {
            b = nbr[(long long)a * K + p];
            if (b >= 0 && garbage && garbage[b]) b = -1;
        }
        // a neighbour listed twice shares the mask copy of its first entry
        int slot = p;
        for (int j = 0; j < n; ++j) {
            const int bj = __shfl(b, j);
            if (j < slot && bj == b && b >= 0) slot = j;
        }
        if (p < n) {
            s_nb[p] = b;
            s_slot[p] = slot;
        }
        if (p == 0) {
            s_any = 0;
            s_chg[0] = 0;
            s_chg[1] = 0;
        }
    }
    const long long base = (long long)a * WW;
    const uint32_t cb = p < WW ? cons[base + p] : 0u;
    const uint32_t nmask = (1u << n) - 1u;
    const bool walk = (cb >> 31) && __popc(cb & nmask) >= 2;
    // the pixel's source in every consistent neighbour, straight into registers
    int ci[kBpMaxNb];
#pragma unroll
    for (int j = 0; j < kBpMaxNb; ++j) ci[j] = 0;
    if (walk) {
        const unsigned short* sp = srcs + (base + p) * K;
#pragma unroll
        for (int j = 0; j < kBpMaxNb; ++j)
            if (j < n) ci[j] = sp[j];
    }
    __syncthreads();
    for (int q = p; q < (1 + n) * MW; q += blockDim.x) {
        const int e = q / MW, t = q - e * MW;
        const int mi = e == 0 ? a : s_nb[e - 1];
        if (mi < 0 || (e > 0 && s_slot[e - 1] != e - 1)) continue;
        ml[q] = masks[(long long)mi * MW + t];
    }
    for (int q = p; q < n * MW * 32; q += blockDim.x) claim[q] = 0u;
    // an inconsistent neighbour has no source from the consistency pass: it points at pixel 0 of
    // its copy, which it neither reads nor marks
#pragma unroll
    for (int j = 0; j < kBpMaxNb; ++j)
        if (j < n) ci[j] = ((cb >> j) & 1u ? ci[j] : 0) + s_slot[j] * MW * 32;
    __syncthreads();
    // the MI's own mask does not change during its walk
    const bool act = walk && !((ml[p >> 5] >> (p & 31)) & 1u);
    uint32_t inc0 = 0;
    if (act) {
#pragma unroll
        for (int j = 0; j < kBpMaxNb; ++j)
            if (j < n) inc0 |= ((cb >> j) & ~(nm[ci[j] >> 5] >> (ci[j] & 31)) & 1u) << j;
#pragma unroll
        for (int j = 0; j < kBpMaxNb; ++j)
            if ((inc0 >> j) & 1u) atomicAdd(&claim[ci[j]], 1u);
    }
    __syncthreads();
    uint32_t cm = 0;
#pragma unroll
    for (int j = 0; j < kBpMaxNb; ++j)
        if (((inc0 >> j) & 1u) && claim[ci[j]] > 1u) cm |= 1u << j;
    if (cm) s_any = 1;
    bool fz = __popc(inc0) >= 2;
    uint32_t inc = inc0;
    __syncthreads();
    if (s_any) {
        for (uint32_t r = 1;; ++r) {
            if (fz) {
                const uint32_t key = (r << 16) | (0xFFFFu - (uint32_t)p);
#pragma unroll
                for (int j = 0; j < kBpMaxNb; ++j)
                    if ((cm >> j) & 1u) atomicMax(&claim[ci[j]], key);
            }
            __syncthreads();
            if (p == 0) s_chg[(r + 1) & 1] = 0;
            uint32_t ni = inc0;
#pragma unroll
            for (int j = 0; j < kBpMaxNb; ++j) {
                if ((cm >> j) & 1u) {
                    const uint32_t v = claim[ci[j]];
                    if ((v >> 16) == r && 0xFFFFu - (v & 0xFFFFu) < (uint32_t)p) ni &= ~(1u << j);
                }
            }
            const bool nf = act && __popc(ni) >= 2;
            if (nf != fz) s_chg[r & 1] = 1;
            fz = nf;
            inc = ni;
            __syncthreads();
            if (!s_chg[r & 1]) break;
        }
    }
    if (fz) {
        incl[base + p] = inc;
#pragma unroll
        for (int j = 0; j < kBpMaxNb; ++j)
            if ((inc >> j) & 1u) atomicOr(&nm[ci[j] >> 5], 1u << (ci[j] & 31));
    }
    __syncthreads();
    for (int q = p + MW; q < (1 + n) * MW; q += blockDim.x) {
        const int e = q / MW, t = q - e * MW;
        const int mi = s_nb[e - 1];
        if (mi < 0 || s_slot[e - 1] != e - 1) continue;
        masks[(long long)mi * MW + t] = ml[q];
    }
}

// Fused point of MI pixel i (incl bits m): the reference's sums in neighbour order, divided by
// (count + 1) as a product with the float reciprocal (cv::Vec3f /= float).
__device__ void bp_fused(const float4* __restrict__ plane, const float2* __restrict__ centers,
                         const uint8_t* __restrict__ bgr, const int* __restrict__ nbr, int K,
                         const BpGeom& g, long long i, uint32_t m, float* P, float* Nr, float* Cl) {
#pragma clang fp contract(off)
    const int W = g.W, WW = W * W;
    const int a = (int)(i / WW), pix = (int)(i - (long long)a * WW);
    const int c = pix % W, r = pix / W;
    const float4 pr = plane[i];
    const float vd_ref = bp_vd(pr.w, g.mla_base);
    const float2 ca = centers[a];
    P[0] = ((float)c - g.cmi) * vd_ref + ca.x;
    P[1] = ((float)r - g.cmi) * vd_ref + ca.y;
    P[2] = vd_ref;
    Nr[0] = pr.x; Nr[1] = pr.y; Nr[2] = pr.z;
    const uint8_t* cr = bgr + i * 3;
    Cl[0] = (float)cr[0]; Cl[1] = (float)cr[1]; Cl[2] = (float)cr[2];
    for (uint32_t mm = m; mm; mm &= mm - 1) {
        const int j = __ffs(mm) - 1;
        const int b = nbr[(long long)a * K + j];
        const float2 cb = centers[b];
        int sr, sc;
        bp_source(vd_ref, c, r, ca, cb, g, sr, sc);
        const long long is = (long long)b * WW + sr * W + sc;
        const float4 ps = plane[is];
        const float vd_src = bp_vd(ps.w, g.mla_base);
        P[0] = P[0] + (((float)sc - g.cmi) * vd_src + cb.x);
        P[1] = P[1] + (((float)sr - g.cmi) * vd_src + cb.y);
        P[2] = P[2] + vd_src;
        Nr[0] = Nr[0] + ps.x; Nr[1] = Nr[1] + ps.y; Nr[2] = Nr[2] + ps.z;
        const uint8_t* cs = bgr + is * 3;
        Cl[0] = Cl[0] + (float)cs[0]; Cl[1] = Cl[1] + (float)cs[1]; Cl[2] = Cl[2] + (float)cs[2];
    }
    const float denom = (float)__popc(m) + 1.0f;
    const float ial = 1.0f / denom;
    for (int k = 0; k < 3; ++k) {
        P[k] = P[k] * ial;
        Nr[k] = Nr[k] * ial;
        Cl[k] = Cl[k] * ial;
    }
}

__device__ __forceinline__ uint32_t sat_u8(float v) {
    const int iv = __float2int_rn(v);  // cvRound: nearest, ties to even
    return (uint32_t)min(max(iv, 0), 255);
}

// Pass 1 (POINTS false): splat position, winner by walk index, in-frame flag.  Pass 2 (POINTS
// true): point records {x, y, z, nx, ny, nz, b, g, r} at their place in the push order.
template <bool POINTS>
__global__ void __launch_bounds__(256) k_bp_emit(
    const float4* __restrict__ plane, const float2* __restrict__ centers,
    const uint8_t* __restrict__ bgr, const int* __restrict__ nbr, int K, BpGeom g, long long M,
    int cols, int rows, const uint32_t* __restrict__ incl, int* __restrict__ pos,
    uint32_t* __restrict__ winner, int* __restrict__ flag, const int* __restrict__ offs,
    float* __restrict__ points, unsigned long long* __restrict__ counts) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    bool fused = false;
    if (i < M) {
        const uint32_t m = incl[i];
        if (m) {
            fused = true;
            float P[3], Nr[3], Cl[3];
            bp_fused(plane, centers, bgr, nbr, K, g, i, m, P, Nr, Cl);
            const int vx = __float2int_rn(P[0]), vy = __float2int_rn(P[1]);
            const bool in = vy >= 0 && vy < rows && vx >= 0 && vx < cols;
            if (!POINTS) {
                pos[i] = in ? vy * cols + vx : -1;
                flag[i] = in ? 1 : 0;
                if (in) atomicMax(&winner[(long long)vy * cols + vx], (uint32_t)(i + 1));
            } else if (in) {
                float* o = points + (long long)offs[i] * 9;
                for (int k = 0; k < 3; ++k) { o[k] = P[k]; o[3 + k] = Nr[k]; o[6 + k] = Cl[k]; }
            }
        } else if (!POINTS) {
            pos[i] = -1;
            flag[i] = 0;
        }
    }
    if (!POINTS) {
        __shared__ int s_n;
        if (threadIdx.x == 0) s_n = 0;
        __syncthreads();
        const unsigned long long b = __ballot(fused);
        if (b && (threadIdx.x & 63) == __ffsll((long long)b) - 1) atomicAdd(&s_n, __popcll(b));
        __syncthreads();
        if (threadIdx.x == 0 && s_n) atomicAdd(&counts[1], (unsigned long long)s_n);
    }
}

// Each splat position takes the fused point with the highest walk index (the reference writes
// them in walk order, the last one stays).
__global__ void __launch_bounds__(256) k_bp_resolve(
    const float4* __restrict__ plane, const float2* __restrict__ centers,
    const uint8_t* __restrict__ bgr, const int* __restrict__ nbr, int K, BpGeom g, long long M,
    const uint32_t* __restrict__ incl, const int* __restrict__ pos,
    const uint32_t* __restrict__ winner, float* __restrict__ vd_out, uint8_t* __restrict__ aif,
    uint8_t* __restrict__ conf) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int o = pos[i];
    if (o < 0 || winner[o] != (uint32_t)(i + 1)) return;
    const uint32_t m = incl[i];
    float P[3], Nr[3], Cl[3];
    bp_fused(plane, centers, bgr, nbr, K, g, i, m, P, Nr, Cl);
    vd_out[o] = P[2];
    conf[o] = (uint8_t)min(255, (__popc(m) + 1) * 60);
    aif[(long long)o * 3] = (uint8_t)sat_u8(Cl[0]);
    aif[(long long)o * 3 + 1] = (uint8_t)sat_u8(Cl[1]);
    aif[(long long)o * 3 + 2] = (uint8_t)sat_u8(Cl[2]);
}

__global__ void k_bp_count(const int* __restrict__ offs, const int* __restrict__ flag, long long M,
                           unsigned long long* __restrict__ counts) {
    counts[0] = (unsigned long long)(M > 0 ? offs[M - 1] + flag[M - 1] : 0);
}

struct BpLayout {
    size_t cons, srcs, masks, incl, pos, flag, offs, sums, winner, total;
};

BpLayout bp_layout(int V, int crop, int raw_w, int raw_h) {
    const size_t M = (size_t)V * crop * crop;
    const size_t MW = (size_t)((crop * crop + 31) / 32);
    const size_t N = (size_t)raw_w * raw_h;
    BpLayout L;
    size_t o = 0;
    L.cons = o; o += al(M * 4);
    L.srcs = o; o += al(M * 31 * 2);
    L.masks = o; o += al((size_t)V * MW * 4);
    L.incl = o; o += al(M * 4);
    L.pos = o; o += al(M * 4);
    L.flag = o; o += al((M + 1) * 4);
    L.offs = o; o += al((M + 1) * 4);
    L.sums = o; o += scan_sums_bytes((long long)M + 1);
    L.winner = o; o += al(N * 4);
    L.total = o;
    return L;
}

int check_common(int V, int crop, int raw_w, int raw_h, const char* what) {
    if (V <= 0 || crop <= 0 || raw_w <= 0 || raw_h <= 0)
        return set_error(LFPM_E_INVALID, "%s: bad sizes (V %d, crop %d, %dx%d)", what, V, crop,
                         raw_w, raw_h);
    if (crop > 127)
        return set_error(LFPM_E_UNSUPPORTED, "%s: crop %d > 127", what, crop);
    if ((long long)V * crop * crop >= (long long)kSeqMask || V >= (1 << 18) - 1)
        return set_error(LFPM_E_UNSUPPORTED, "%s: %d MIs of %d^2 pixels exceed the index range",
                         what, V, crop);
    if ((long long)raw_w * raw_h >= INT_MAX)
        return set_error(LFPM_E_UNSUPPORTED, "%s: frame %dx%d too large", what, raw_w, raw_h);
    return LFPM_OK;
}

}  // namespace
}  // namespace lfpm

using namespace lfpm;

extern "C" {

size_t lfpm_aif_workspace_size(int num_mis, int crop, int raw_w, int raw_h) {
    (void)num_mis; (void)crop;
    return aif_layout(raw_w, raw_h).total;
}

int lfpm_aif_composite(const float* plane_dev, const float* centers_dev,
                       const uint8_t* bgr_tiles_dev, const uint8_t* richness_tiles_dev,
                       const uint8_t* blur_tiles_dev, const uint8_t* garbage_dev, int num_mis,
                       int crop, float baseline, float B, float bl0, float main_lens_f,
                       int raw_w, int raw_h, float* vd_dev, float* inv_vd_dev,
                       float* real_mm_dev, uint8_t* aif_dev, uint8_t* filled_dev,
                       uint8_t* denoised_dev, int* stats_dev, void* workspace_dev,
                       size_t workspace_bytes, void* stream) {
    int rc = check_common(num_mis, crop, raw_w, raw_h, "lfpm_aif_composite");
    if (rc != LFPM_OK) return rc;
    if (!plane_dev || !centers_dev || !bgr_tiles_dev || !vd_dev || !inv_vd_dev || !real_mm_dev ||
        !aif_dev || !filled_dev || !denoised_dev || !stats_dev || !workspace_dev)
        return set_error(LFPM_E_INVALID, "lfpm_aif_composite: null pointer");
    const AifLayout L = aif_layout(raw_w, raw_h);
    if (workspace_bytes < L.total)
        return set_error(LFPM_E_INVALID, "lfpm_aif_composite: workspace %zu < %zu bytes",
                         workspace_bytes, L.total);
    const size_t lds_rows = (size_t)raw_w * 8;
    if (lds_rows > 65536)
        return set_error(LFPM_E_UNSUPPORTED, "lfpm_aif_composite: frame width %d > 8192",
                         raw_w);
    const int stage_kcrop = lds_rows + (size_t)crop * crop * 5 <= 65536 ? 1 : 0;
    hipStream_t s = (hipStream_t)stream;
    unsigned char* ws = (unsigned char*)workspace_dev;
    uint32_t* win = (uint32_t*)(ws + L.win);
    int* cand = (int*)(ws + L.cand);
    uint32_t* col32 = (uint32_t*)(ws + L.col32);
    uint32_t* map32 = (uint32_t*)(ws + L.map32);
    uint8_t* planes = ws + L.planes;
    const long long N = (long long)raw_w * raw_h;
    const long long M = (long long)num_mis * crop * crop;
    k_fill_u32<<<grid_for(N, 256), 256, 0, s>>>(win, N, 0u);
    k_fill_u32<<<grid_for(N, 256), 256, 0, s>>>((uint32_t*)cand, N, 0u);
    k_fill_u32<<<1, 64, 0, s>>>((uint32_t*)stats_dev, 8, 0u);
    k_fill_u32<<<1, 64, 0, s>>>((uint32_t*)stats_dev + 5, 1, 0xFFFFFFFFu);
    k_aif_collect<<<grid_for(M, 256), 256, 0, s>>>(
        (const float4*)plane_dev, (const float2*)centers_dev, richness_tiles_dev, blur_tiles_dev,
        garbage_dev, num_mis, crop, baseline, raw_w, raw_h, win, cand, stats_dev);
    k_aif_resolve<<<grid_for(N, 256), 256, 0, s>>>(
        (const float4*)plane_dev, bgr_tiles_dev, win, cand, crop, baseline, B, bl0, main_lens_f,
        N, vd_dev, inv_vd_dev, real_mm_dev, aif_dev, col32, map32, stats_dev);
    // HoleFilling: passes of row fixed-point iterations; the stream is synchronised to read the
    // iterations' dirty counts and the pass state
    const size_t lds = lds_rows + (stage_kcrop ? (size_t)crop * crop * 5 : 0);
    HoleState h;
    h.a_col = col32; h.a_map = map32;
    h.b_col0 = (uint32_t*)(ws + L.b[0]); h.b_map0 = (uint32_t*)(ws + L.b[1]);
    h.b_col1 = (uint32_t*)(ws + L.b[2]); h.b_map1 = (uint32_t*)(ws + L.b[3]);
    h.slot = (int*)(ws + L.slot); h.ver = (int*)(ws + L.ver); h.dep = (int*)(ws + L.dep);
    h.fills = (int*)(ws + L.fills); h.dirty = (int*)(ws + L.dirty);
    int* pstate = (int*)(ws + L.pstate);
    k_fill_u32<<<1, 64, 0, s>>>((uint32_t*)pstate, 4, 0u);
    k_hole_count<<<grid_for(N, 256), 256, 0, s>>>(col32, N, pstate);
    k_aif_kstar<<<1, 1, 0, s>>>(stats_dev, crop * crop, pstate);
    const int niter = raw_h + 2 + kHoleGroup;
    std::vector<int> dirty(kHoleGroup);
    int pst[4] = {0, 0, 0, 0};
    for (;;) {
        k_hole_init<<<1, 256, 0, s>>>(h.slot, h.ver, raw_h, h.dirty, niter);
        int t = 1, fin = -1;
        while (fin < 0) {
            for (int g = 0; g < kHoleGroup; ++g)
                k_hole_row<<<raw_h, kRowThreads, lds, s>>>(h, bgr_tiles_dev, crop, raw_w, raw_h,
                                                           stats_dev, stage_kcrop, t + g);
            if (hipMemcpyAsync(dirty.data(), h.dirty + t, kHoleGroup * sizeof(int),
                               hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                return check_launch("lfpm_aif_composite (hole filling)");
            for (int g = 0; g < kHoleGroup && fin < 0; ++g)
                if (dirty[g] == 0) fin = t + g;
            t += kHoleGroup;
            if (fin < 0 && t > raw_h + 1)
                return set_error(LFPM_E_HIP, "lfpm_aif_composite: hole filling did not settle");
        }
        k_hole_gather<<<grid_for(N, 256), 256, 0, s>>>(h, raw_w, raw_h, fin & 1, col32, map32);
        k_hole_pass_end<<<1, 256, 0, s>>>(h.fills, raw_h, raw_w, pstate, stats_dev);
        if (hipMemcpyAsync(pst, pstate, sizeof(pst), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return check_launch("lfpm_aif_composite (hole filling)");
        if (pst[3]) break;
    }
    k_unpack_rgb<<<grid_for(N, 256), 256, 0, s>>>(col32, N, filled_dev);
    // denoise: five 5x5 median passes of the composite, per channel
    k_split3<<<grid_for(N, 256), 256, 0, s>>>(aif_dev, N, planes);
    for (int c = 0; c < 3; ++c)
        median5_passes(planes + c * N, ws + L.tmp0, ws + L.tmp1, raw_w, raw_h, 5,
                       denoised_dev + c, 3, stream);
    return check_launch("lfpm_aif_composite");
}

int lfpm_fuse_schedule(const int32_t* nbr, const int32_t* cnt, const uint8_t* garbage,
                       int num_mis, int max_neighbors, int32_t* order, int32_t* level_start,
                       int* num_levels) {
    if (!nbr || !cnt || !order || !level_start || !num_levels || num_mis <= 0 ||
        max_neighbors <= 0)
        return set_error(LFPM_E_INVALID, "lfpm_fuse_schedule: bad arguments");
    const int V = num_mis, K = max_neighbors;
    auto live = [&](int b) { return b >= 0 && b < V && !(garbage && garbage[b]); };
    std::vector<std::vector<int>> inv(V);
    for (int a = 0; a < V; ++a) {
        if (!live(a)) continue;
        const int n = std::min(std::max(cnt[a], 0), K);
        for (int j = 0; j < n; ++j) {
            const int b = nbr[(size_t)a * K + j];
            if (live(b)) inv[b].push_back(a);
        }
    }
    // level(b) = 1 + the largest level of an earlier MI whose walk touches a mask b's walk
    // touches: a writes the masks of its neighbours N(a) and reads them and its own
    std::vector<int> level(V, -1);
    int nl = 0;
    for (int b = 0; b < V; ++b) {
        if (!live(b)) continue;
        int lv = 0;
        auto dep = [&](int a) {
            if (a < b && level[a] >= 0) lv = std::max(lv, level[a] + 1);
        };
        const int n = std::min(std::max(cnt[b], 0), K);
        for (int a : inv[b]) dep(a);  // a's neighbours include b
        for (int j = 0; j < n; ++j) {
            const int m = nbr[(size_t)b * K + j];
            if (!live(m)) continue;
            dep(m);                      // b's neighbours include m
            for (int a : inv[m]) dep(a);  // a and b share the neighbour m
        }
        level[b] = lv;
        nl = std::max(nl, lv + 1);
    }
    std::vector<int> count(nl + 1, 0);
    for (int b = 0; b < V; ++b)
        if (level[b] >= 0) ++count[level[b] + 1];
    for (int l = 0; l < nl; ++l) count[l + 1] += count[l];
    for (int l = 0; l <= nl; ++l) level_start[l] = count[l];
    std::vector<int> fillp(count.begin(), count.end() - 1);
    for (int b = 0; b < V; ++b)
        if (level[b] >= 0) order[fillp[level[b]]++] = b;
    *num_levels = nl;
    return LFPM_OK;
}

size_t lfpm_fuse_workspace_size(int num_mis, int crop, int raw_w, int raw_h) {
    return bp_layout(num_mis, crop, raw_w, raw_h).total;
}

int lfpm_fuse_backproject(const float* plane_dev, const float* centers_dev,
                          const uint8_t* bgr_tiles_dev, const int32_t* nbr_dev,
                          const int32_t* cnt_dev, int max_neighbors, const uint8_t* garbage_dev,
                          const int32_t* order_dev, const int32_t* level_start, int num_levels,
                          int num_mis, int crop, float baseline, float B, float bl0, int raw_w,
                          int raw_h, float* vd_dev, uint8_t* aif_dev, uint8_t* conf_dev,
                          float* points_dev, size_t points_bytes,
                          unsigned long long* counts_dev, void* workspace_dev,
                          size_t workspace_bytes, void* stream) {
    int rc = check_common(num_mis, crop, raw_w, raw_h, "lfpm_fuse_backproject");
    if (rc != LFPM_OK) return rc;
    if (!plane_dev || !centers_dev || !bgr_tiles_dev || !nbr_dev || !cnt_dev || !order_dev ||
        !level_start || !vd_dev || !aif_dev || !conf_dev || !counts_dev || !workspace_dev)
        return set_error(LFPM_E_INVALID, "lfpm_fuse_backproject: null pointer");
    if (max_neighbors <= 0 || max_neighbors > 31)
        return set_error(LFPM_E_UNSUPPORTED,
                         "lfpm_fuse_backproject: max_neighbors %d outside 1..31", max_neighbors);
    if (num_levels < 0 || num_levels > num_mis)
        return set_error(LFPM_E_INVALID, "lfpm_fuse_backproject: %d levels", num_levels);
    const long long M = (long long)num_mis * crop * crop;
    if (points_dev && points_bytes < (size_t)M * 9 * sizeof(float))
        return set_error(LFPM_E_INVALID, "lfpm_fuse_backproject: points buffer %zu < %lld bytes",
                         points_bytes, M * 9 * (long long)sizeof(float));
    const BpLayout L = bp_layout(num_mis, crop, raw_w, raw_h);
    if (workspace_bytes < L.total)
        return set_error(LFPM_E_INVALID, "lfpm_fuse_backproject: workspace %zu < %zu bytes",
                         workspace_bytes, L.total);
    const int WW = crop * crop, MW = (WW + 31) / 32;
    // the workgroup walk up to 1024 pixels per MI, the wavefront walk beyond
    const size_t lds_wg = (size_t)((1 + max_neighbors) * MW + max_neighbors * MW * 32) * 4;
    const bool wg_walk = WW <= kBpMaxPix && lds_wg <= 65536;
    const size_t lds = (size_t)((1 + max_neighbors) * MW + 2 * WW) * 4 +
                       (size_t)WW * max_neighbors * 2;
    if (!wg_walk && lds > 65536)
        return set_error(LFPM_E_UNSUPPORTED, "lfpm_fuse_backproject: %zu B of LDS", lds);
    hipStream_t s = (hipStream_t)stream;
    unsigned char* ws = (unsigned char*)workspace_dev;
    uint32_t* cons = (uint32_t*)(ws + L.cons);
    unsigned short* srcs = (unsigned short*)(ws + L.srcs);
    uint32_t* masks = (uint32_t*)(ws + L.masks);
    uint32_t* incl = (uint32_t*)(ws + L.incl);
    int* pos = (int*)(ws + L.pos);
    int* flag = (int*)(ws + L.flag);
    int* offs = (int*)(ws + L.offs);
    int* sums = (int*)(ws + L.sums);
    uint32_t* winner = (uint32_t*)(ws + L.winner);
    const long long N = (long long)raw_w * raw_h;
    BpGeom g;
    g.mla_base = baseline * bl0 / (bl0 + B);
    g.cmi = (float)(crop - 1) * 0.5f + 1.0f;
    g.W = crop;
    const float4* pl = (const float4*)plane_dev;
    const float2* ce = (const float2*)centers_dev;
    k_fill_u32<<<grid_for((long long)num_mis * MW, 256), 256, 0, s>>>(masks,
                                                                      (long long)num_mis * MW, 0u);
    k_fill_u32<<<grid_for(M, 256), 256, 0, s>>>(incl, M, 0u);
    k_fill_u32<<<grid_for(N, 256), 256, 0, s>>>(winner, N, 0u);
    k_fill_u32<<<grid_for(N, 256), 256, 0, s>>>((uint32_t*)vd_dev, N, 0u);
    k_fill_u8<<<grid_for(3 * N, 256), 256, 0, s>>>(aif_dev, 3 * N);
    k_fill_u8<<<grid_for(N, 256), 256, 0, s>>>(conf_dev, N);
    k_fill_u32<<<1, 64, 0, s>>>((uint32_t*)counts_dev, 4, 0u);
    k_fill_u32<<<1, 64, 0, s>>>((uint32_t*)(flag + M), 1, 0u);
    k_bp_consistency<<<grid_for(M, 256), 256, 0, s>>>(pl, ce, nbr_dev, cnt_dev, max_neighbors,
                                                      garbage_dev, num_mis, g, cons, srcs);
    for (int l = 0; l < num_levels; ++l) {
        const int b0 = level_start[l], nb = level_start[l + 1] - b0;
        if (nb <= 0) continue;
        if (wg_walk)
            k_bp_level_wg<<<nb, (WW + 63) / 64 * 64, lds_wg, s>>>(
                order_dev, b0, nbr_dev, cnt_dev, max_neighbors, garbage_dev, crop, cons, srcs,
                masks, incl);
        else
            k_bp_level<<<nb, 64, lds, s>>>(order_dev, b0, pl, ce, nbr_dev, cnt_dev,
                                           max_neighbors, garbage_dev, g, cons, srcs, masks, incl);
    }
    k_bp_emit<false><<<grid_for(M, 256), 256, 0, s>>>(pl, ce, bgr_tiles_dev, nbr_dev,
                                                     max_neighbors, g, M, raw_w, raw_h, incl, pos,
                                                     winner, flag, nullptr, nullptr, counts_dev);
    k_bp_resolve<<<grid_for(M, 256), 256, 0, s>>>(pl, ce, bgr_tiles_dev, nbr_dev, max_neighbors,
                                                  g, M, incl, pos, winner, vd_dev, aif_dev,
                                                  conf_dev);
    exclusive_scan(flag, M + 1, offs, sums, stream);
    k_bp_count<<<1, 1, 0, s>>>(offs, flag, M, counts_dev);
    if (points_dev)
        k_bp_emit<true><<<grid_for(M, 256), 256, 0, s>>>(pl, ce, bgr_tiles_dev, nbr_dev,
                                                        max_neighbors, g, M, raw_w, raw_h, incl,
                                                        pos, winner, flag, offs, points_dev,
                                                        counts_dev);
    return check_launch("lfpm_fuse_backproject");
}

}  // extern "C"
