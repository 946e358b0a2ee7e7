// lfpm_kernels.hip — hand-written HIP kernels (gfx950 / CDNA4) of the micro-image PatchMatch
// solver and the C-ABI entry points that launch them.
//
// Work decomposition.  One workgroup owns one micro-image (MI) at a time and loops over MIs
// (grid = resident slots, each XCD walks a contiguous MI range so ring-1/ring-2 neighbour
// tiles stay in that XCD's L2).  The MI's reference tile and its K neighbour tiles are staged
// in LDS; one lane owns one pixel of the active red/black colour and evaluates all of that
// pixel's hypotheses; every bilinear gather of the bilateral NCC is an LDS read.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/lfpm.h"
#include "lfpm_common.h"
#include "lfpm_internal.h"

namespace lfpm {

// ------------------------------------------------------------------------------------------
// Parameters as the kernels see them
// ------------------------------------------------------------------------------------------
struct DevParams {
    int W, WW, K, r, inc, top_k, V;
    int W1, TS;          // padded tile row stride (W+1) and slot stride (W+1)^2 in LDS
    int tile_bytes;      // bytes per tile cell in LDS (by the tile format)
    int fmt;             // tile format: 0 FP32 values, 3 FP16 code pairs (storage 1)
    float base;          // Base (disparity unit, lens pitch in px)
    float inv_base;      // 1 / max(Base, 1e-6)
    float inv_ss;        // 1 / (2 sigma_s^2)
    float inv_sc;        // 1 / (2 sigma_c^2)
    float min_var;
    float dmin, dmax;
    float cos_max_slant;
    uint32_t seed, uid;
    int slots_per_xcd;   // persistent grid: workgroups per XCD
    int blur;            // blur term on (blur_lambda > 0; wave engine, FP32 tiles)
    float blur_lambda;
};

static DevParams make_dev_params(const lfpm_pm_params* p, int V) {
    DevParams d;
    d.W = p->crop;
    d.WW = p->crop * p->crop;
    d.W1 = p->crop + 1;
    d.TS = (p->crop + 1) * (p->crop + 1);
    d.fmt = p->storage == 1 ? 3 : 0;
    d.tile_bytes = 4;
    d.K = p->max_neighbors;
    d.r = p->radius;
    d.inc = std::max(1, p->radius_increment);
    d.top_k = p->top_k;
    d.V = V;
    d.base = p->base;
    d.inv_base = 1.0f / std::max(p->base, 1e-6f);
    d.inv_ss = 1.0f / (2.0f * p->sigma_spatial * p->sigma_spatial);
    d.inv_sc = 1.0f / (2.0f * p->sigma_color * p->sigma_color);
    d.min_var = p->min_var;
    d.dmin = p->depth_min;
    d.dmax = p->depth_max;
    d.cos_max_slant = p->cos_max_slant;
    d.seed = p->seed;
    d.uid = p->frame_uid;
    d.slots_per_xcd = 1;
    d.blur = p->blur_lambda > 0.0f ? 1 : 0;
    d.blur_lambda = p->blur_lambda;
    return d;
}

static int check_pm_params(const lfpm_pm_params* p, int V) {
    if (!p) return set_error(LFPM_E_INVALID, "null params");
    if (V <= 0) return set_error(LFPM_E_INVALID, "num_mis must be > 0");
    if (p->crop < 3 || p->crop > LFPM_MAX_CROP)
        return set_error(LFPM_E_INVALID, "crop %d outside [3, %d]", p->crop, LFPM_MAX_CROP);
    if (p->max_neighbors < 1 || p->max_neighbors > LFPM_MAX_NEIGHBORS)
        return set_error(LFPM_E_INVALID, "max_neighbors %d outside [1, %d]", p->max_neighbors,
                         LFPM_MAX_NEIGHBORS);
    if (p->radius < 0 || p->radius > LFPM_MAX_RADIUS)
        return set_error(LFPM_E_INVALID, "radius %d outside [0, %d]", p->radius, LFPM_MAX_RADIUS);
    if (p->top_k < 1) return set_error(LFPM_E_INVALID, "top_k must be >= 1");
    if (!(p->sigma_spatial > 0.0f) || !(p->sigma_color > 0.0f))
        return set_error(LFPM_E_INVALID, "sigmas must be > 0");
    if (p->storage != 0 && p->storage != 1)
        return set_error(LFPM_E_UNSUPPORTED, "storage mode %d", p->storage);
    if (!(p->blur_lambda >= 0.0f) || !std::isfinite(p->blur_lambda))
        return set_error(LFPM_E_INVALID, "blur_lambda must be finite and >= 0");
    if (p->blur_lambda > 0.0f && p->storage != 0)
        return set_error(LFPM_E_UNSUPPORTED, "blur term: storage must be 0 (FP32 tiles)");
    return LFPM_OK;
}

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return set_error(LFPM_E_HIP, "%s: %s", #expr, hipGetErrorString(e_));        \
    } while (0)

static int launch_check(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(LFPM_E_HIP, "%s launch: %s", what, hipGetErrorString(e));
    return LFPM_OK;
}

// ------------------------------------------------------------------------------------------
// Device helpers
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float u01_dev(uint32_t x) {
    return ((float)(x >> 8) + 0.5f) * 5.9604644775390625e-08f;
}

// Per-(pixel, step) Philox draw stream with a one-block cache.
// x / -0.18f correctly rounded: Markstein's correction of x * RN(1 / -0.18f).  The arguments here
// are squared costs below 0.64 (0, or at least 2^-48); for every float in {0} u [2^-60, 1] the
// result equals the IEEE quotient (checked exhaustively).
__device__ __forceinline__ float div_m018(float x) {
    const float c = -0.18f, y = -0x1.638e38p+2f;  // y = RN(1 / c)
    const float q0 = x * y;
    return fmaf(fmaf(-q0, c, x), y, q0);
}

struct DevStream {
    uint32_t pix, step, k0, k1, blk;
    uint32_t w[4];
    __device__ DevStream(uint32_t p, uint32_t s, uint32_t a, uint32_t b)
        : pix(p), step(s), k0(a), k1(b), blk(0xFFFFFFFFu) {}
    __device__ uint32_t raw(uint32_t slot) {
        const uint32_t b = slot >> 2;
        if (b != blk) {
            philox4x32_10(pix, step, b, 0u, k0, k1, w);
            blk = b;
        }
        const uint32_t q = slot & 3u;
        return q == 0 ? w[0] : (q == 1 ? w[1] : (q == 2 ? w[2] : w[3]));
    }
    // Slot slot0 + l, where every lane's slot lies in slot0's block: the block is computed once
    // for the wavefront (slot0 is uniform) and each lane picks its word.
    __device__ uint32_t raw_lane(uint32_t slot0, uint32_t l) {
        (void)raw(slot0);
        const uint32_t q = (slot0 + l) & 3u;
        return q == 0 ? w[0] : (q == 1 ? w[1] : (q == 2 ? w[2] : w[3]));
    }
};

__device__ __forceinline__ void normalize3(float& x, float& y, float& z) {
    const float n = sqrtf(x * x + y * y + z * z);
    x = x / n; y = y / n; z = z / n;
}

// Marsaglia (1972) sphere point; rejection decided exactly in 64-bit integers.
__device__ void random_normal(DevStream& st, uint32_t base, float& nx, float& ny, float& nz) {
    long long A = 0, B = 0;
    for (uint32_t a = 0; a < MARSAGLIA_ATTEMPTS; ++a) {
        const long long ka = (long long)(st.raw(base + 2 * a) >> 8);
        const long long kb = (long long)(st.raw(base + 2 * a + 1) >> 8);
        const long long ta = 2 * ka + 1 - (1LL << 24);
        const long long tb = 2 * kb + 1 - (1LL << 24);
        if (ta * ta + tb * tb < (1LL << 48)) { A = ta; B = tb; break; }
    }
    const float q1 = (float)A * 5.9604644775390625e-08f, q2 = (float)B * 5.9604644775390625e-08f;
    const float s = q1 * q1 + q2 * q2;
    const float sq = sqrtf(1.0f - s);
    nx = 2.0f * q1 * sq;
    ny = 2.0f * q2 * sq;
    nz = 1.0f - 2.0f * s;
    normalize3(nx, ny, nz);
}

// Bilateral NCC between the reference tile and one neighbour tile (both in LDS).
// Follows ComputeBilateralNCC_Frame (MIStereo/AdaptMIPMFrame.cu:73-179): window (2r+1)^2 with
// stride inc, samples whose reference or source point leaves the crop are skipped, weights
// exp(-|dxy|/(2 s_s^2) - |dI|/(2 s_c^2)), moments taken about the window centre intensity.
// Work counters of the counting kernel variants: evaluated window samples and evaluated
// (pixel, view, hypothesis) triples — the units of the roofline FLOP count.
struct WorkCount {
    unsigned int samples, triples;
    // the samples and triples a single-round scoring would add (rows settled by the bound)
    unsigned int samples_skipped, triples_skipped;
};

// LDS carve-up shared by the solver kernels.
struct Lds {
    float* tiles;     // (K+1) slots of (W+1)^2 cells: slot 0 = reference, slot 1+k = neighbour
                      // k; column W and row W repeat the last column/row (clamp-to-edge for
                      // free); FP32 values or PairH cells
    float* costp;     // WW (phase kernel: previous cost of this MI)
    float* spat;      // (2r+1)^2 spatial log-weights -|dxy| / (2 s_s^2)
    float2* step;     // K: (c_ref - c_nei) / Base
    float2* cdiff;    // K: c_ref - c_nei
    int* nid;         // K
    float* btiles;    // blur term: blur values of the same (K+1) slots, FP32 (else null)
    unsigned* next;   // wave engine: next pixel of the current MI to hand out
    int* cur;         // wave engine: the workgroup's current MI (-1: none left)
};

// Tile formats in LDS; arithmetic is FP32 in both.
//  float  (storage 0): the FP32 tiles as given.
//  PairH  (storage 1): the 8-bit intensity codes 0..255 (k/255 in the FP32 tiles, recovered
//         exactly by rint(255 t)) in FP16, each cell holding its code t(x, y) and the step to the
//         right d = t(x+1, y) - t(x, y) (0 in the last column, whose right neighbour is itself
//         under clamp-to-edge).  Integers up to 2048 are exact in FP16, so the storage loses
//         nothing; a row's two bilinear taps are one read and one FMA, t + fx * d, the rounding
//         of the two-tap form t(x, y) + fx * (t(x+1, y) - t(x, y)).  The NCC then runs on the
//         0..255 scale: it is scale invariant, and the two scale-dependent constants (colour
//         weight 1/(2 s_c^2) and the variance gate) are rescaled by TileScale.
struct __align__(4) PairH { __half t, d; };
template <typename TT> struct IsPair { static constexpr bool value = false; };
template <> struct IsPair<PairH> { static constexpr bool value = true; };
template <typename TT> __device__ __forceinline__ const TT* tile_base(const Lds& L) {
    return (const TT*)L.tiles;
}
__device__ __forceinline__ float ldt(const float* p) { return *p; }
__device__ __forceinline__ float ldt(const PairH* p) { return __half2float(p->t); }
__device__ __forceinline__ void stt(float* p, float v) { *p = v; }
__device__ __forceinline__ void stt_pair(PairH* p, float v, float vr) {
    const float c = rintf(v * 255.0f), cr = rintf(vr * 255.0f);
    *p = PairH{__float2half_rn(c), __float2half_rn(cr - c)};
}
__device__ __forceinline__ float pair_t(const PairH& p) { return __half2float(p.t); }
__device__ __forceinline__ float pair_d(const PairH& p) { return __half2float(p.d); }
template <typename TT> struct TileScale;
template <> struct TileScale<float> {
    __device__ static float inv_sc(const DevParams& P) { return P.inv_sc; }
    __device__ static float min_var(const DevParams& P) { return P.min_var; }
};
template <> struct TileScale<PairH> {
    __device__ static float inv_sc(const DevParams& P) { return P.inv_sc * (1.0f / 255.0f); }
    __device__ static float min_var(const DevParams& P) { return P.min_var * 65025.0f; }
};

__device__ __forceinline__ void fill_spatial(Lds& L, const DevParams& P) {
    const int n = 2 * P.r + 1;
    for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) {
        const int j = idx / n - P.r, i = idx % n - P.r;
        L.spat[idx] = -sqrtf((float)(i * i + j * j)) * P.inv_ss;
    }
}

// Stage the MI's tables and tiles (reference + its neighbours) into LDS.  Caller syncs.
template <typename TT>
__device__ __forceinline__ int stage_mi(Lds& L, const DevParams& P, int mi,
                                        const float* __restrict__ tiles,
                                        const float2* __restrict__ centers,
                                        const int* __restrict__ nbr, const int* __restrict__ cnt,
                                        const float* __restrict__ btiles = nullptr) {
    const int K = P.K, WW = P.WW;
    const int c = min(cnt[mi], K);
    const float2 cr = centers[mi];
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        const int n = k < c ? nbr[(size_t)mi * K + k] : -1;
        L.nid[k] = n;
        if (n >= 0) {
            const float2 cn = centers[n];
            const float dx = cr.x - cn.x, dy = cr.y - cn.y;
            L.cdiff[k] = make_float2(dx, dy);
            L.step[k] = make_float2(dx * P.inv_base, dy * P.inv_base);
        } else {
            L.cdiff[k] = make_float2(0.0f, 0.0f);
            L.step[k] = make_float2(0.0f, 0.0f);
        }
    }
    const int W = P.W, W1 = P.W1, TS = P.TS;
    const int total = (c + 1) * TS;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int s = idx / TS, o = idx - s * TS;
        const int yy = min(o / W1, W - 1), xx = min(o - (o / W1) * W1, W - 1);
        int src = mi;
        if (s > 0) {
            src = nbr[(size_t)mi * K + s - 1];
            if (src < 0) src = mi;  // never read: views with nid < 0 are not scored
        }
        if constexpr (IsPair<TT>::value)
            stt_pair((TT*)tile_base<TT>(L) + idx, tiles[(size_t)src * WW + yy * W + xx],
                     tiles[(size_t)src * WW + yy * W + min(xx + 1, W - 1)]);
        else
            stt((TT*)tile_base<TT>(L) + idx, tiles[(size_t)src * WW + yy * W + xx]);
        if (btiles) L.btiles[idx] = btiles[(size_t)src * WW + yy * W + xx];
    }
    return c;
}

__device__ __forceinline__ float4 disp_record(const Lds& L, const DevParams& P, int k, float w) {
    const float2 cd = L.cdiff[k];
    const float b = sqrtf(cd.x * cd.x + cd.y * cd.y);
    const float dr = fabsf(w) * b * P.inv_base;
    return make_float4(w, dr, b, b / (dr + 1e-7f));
}

template <bool COUNT>
__device__ __forceinline__ void flush_counts(unsigned long long* counters, const WorkCount& wc) {
    if (COUNT && counters) {
        atomicAdd(&counters[0], (unsigned long long)wc.samples);
        atomicAdd(&counters[1], (unsigned long long)wc.triples);
        atomicAdd(&counters[2], (unsigned long long)wc.samples_skipped);
        atomicAdd(&counters[3], (unsigned long long)wc.triples_skipped);
    }
}

// MI assignment of the persistent grid: workgroup g sits on XCD g % 8; XCD x walks the
// contiguous MI range [x*V/8, (x+1)*V/8) with stride slots_per_xcd.
__device__ __forceinline__ void mi_range(const DevParams& P, int& first, int& end, int& stride) {
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    const int lo = (int)(((long long)P.V * xcd) >> 3);
    const int hi = (int)(((long long)P.V * (xcd + 1)) >> 3);
    first = lo + local;
    end = hi;
    stride = P.slots_per_xcd;
}

// ------------------------------------------------------------------------------------------
// Tile gather
// ------------------------------------------------------------------------------------------
__global__ void k_slice(const uint8_t* __restrict__ raw, size_t pitch, const int2* __restrict__ org,
                        int V, int W, float scale_div, float* __restrict__ out) {
    const size_t WW = (size_t)W * W;
    const size_t total = (size_t)V * WW;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        const int v = (int)(i / WW);
        const int o = (int)(i - (size_t)v * WW);
        const int y = o / W, x = o - y * W;
        const int2 g = org[v];
        const float u = (float)raw[(size_t)(g.y + y) * pitch + (size_t)(g.x + x)];
        out[i] = scale_div > 0.0f ? u / scale_div : u;
    }
}

// Colour crops: pixel (x, y) of MI v = bytes B, G, R of the colour frame at its origin + (x, y)
// (problem.m_Image_rgb of DepthSolver::Slice_RawMLAImage, DepthSolver.cpp:1951-1997).
__global__ void k_slice_bgr(const uint8_t* __restrict__ bgr, size_t pitch,
                            const int2* __restrict__ org, int V, int W, uint8_t* __restrict__ out) {
    const size_t WW = (size_t)W * W;
    const size_t total = (size_t)V * WW;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        const int v = (int)(i / WW);
        const int o = (int)(i - (size_t)v * WW);
        const int y = o / W, x = o - y * W;
        const int2 g = org[v];
        const uint8_t* src = bgr + (size_t)(g.y + y) * pitch + (size_t)(g.x + x) * 3;
        out[3 * i] = src[0];
        out[3 * i + 1] = src[1];
        out[3 * i + 2] = src[2];
    }
}

// Laplacian blur score of the gray frame (ImageQualityEstimate::QuantizeBlurByLaplacian,
// Util/ImageQualityEstimate.cpp:211-230: cv::Laplacian, ksize 1 = the 4-neighbour stencil, FP32,
// BORDER_REFLECT_101), gathered straight into the MIs' blur tiles (the crops of
// DepthSolver::SliceAndIndicatlizeForMI, DepthSolver.cpp:2709-2745) and divided by 255 as
// ComputeBilateralNCC_Frame reads it (MIStereo/AdaptMIPMFrame.cu:140-141).
__device__ __forceinline__ int reflect101(int i, int n) {
    if (n == 1) return 0;
    if (i < 0) return -i;
    if (i >= n) return 2 * n - 2 - i;
    return i;
}

__global__ void k_slice_blur(const uint8_t* __restrict__ raw, size_t pitch, int w, int h,
                             const int2* __restrict__ org, int V, int W, float* __restrict__ out) {
    const size_t WW = (size_t)W * W;
    const size_t total = (size_t)V * WW;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        const int v = (int)(i / WW);
        const int o = (int)(i - (size_t)v * WW);
        const int y = o / W, x = o - y * W;
        const int2 g = org[v];
        const int gx = g.x + x, gy = g.y + y;
        auto at = [&](int xx, int yy) {
            return (int)raw[(size_t)reflect101(yy, h) * pitch + (size_t)reflect101(xx, w)];
        };
        const int lap = at(gx - 1, gy) + at(gx + 1, gy) + at(gx, gy - 1) + at(gx, gy + 1) -
                        4 * at(gx, gy);
        out[i] = (float)lap / 255.0f;
    }
}

// BGR -> gray of cv::cvtColor(..., cv::COLOR_BGR2GRAY) on 8-bit images, the conversion of each
// MI crop cut from the colour raw frame (DepthSolver::Slice_RawMLAImage, DepthSolver.cpp:
// 1985-1990; the frame is read with cv::imread(path, 1), :2503).  The map is per pixel, so the
// frame is converted once and then sliced.  OpenCV's 8-bit form of 0.114 B + 0.587 G + 0.299 R
// with 14-bit fixed-point coefficients, rounded half up: Y = (1868 B + 9617 G + 4899 R + 2^13)
// >> 14 (the coefficients sum to 2^14, so a gray pixel keeps its value).
__device__ __forceinline__ uint32_t bgr_luma(uint32_t b, uint32_t g, uint32_t r) {
    return (b * 1868u + g * 9617u + r * 4899u + 8192u) >> 14;
}

// One thread per 16 pixels of a row: three 16-byte loads and one 16-byte store when both row
// slices are 16-byte aligned and the 16 pixels lie inside the row, byte accesses otherwise (row
// tails, pitches that are not multiples of 16).
__global__ void k_bgr_gray(const uint8_t* __restrict__ bgr, size_t bpitch, int w, int h,
                           uint8_t* __restrict__ gray, size_t gpitch) {
    const size_t chunks = (size_t)((w + 15) >> 4);
    const size_t total = chunks * (size_t)h;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < total;
         c += (size_t)gridDim.x * blockDim.x) {
        const size_t y = c / chunks;
        const int x0 = (int)(c - y * chunks) << 4;
        const uint8_t* src = bgr + y * bpitch + (size_t)x0 * 3;
        uint8_t* dst = gray + y * gpitch + (size_t)x0;
        if (x0 + 16 <= w && (((uintptr_t)src | (uintptr_t)dst) & 15) == 0) {
            const uint4* s4 = (const uint4*)src;
            const uint4 q0 = s4[0], q1 = s4[1], q2 = s4[2];
            const uint32_t wd[12] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y,
                                     q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
            uint32_t o[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int p = 0; p < 16; ++p) {
                const int k = 3 * p;
                const uint32_t b = (wd[k >> 2] >> ((k & 3) * 8)) & 255u;
                const uint32_t g = (wd[(k + 1) >> 2] >> (((k + 1) & 3) * 8)) & 255u;
                const uint32_t r = (wd[(k + 2) >> 2] >> (((k + 2) & 3) * 8)) & 255u;
                o[p >> 2] |= bgr_luma(b, g, r) << ((p & 3) * 8);
            }
            *(uint4*)dst = make_uint4(o[0], o[1], o[2], o[3]);
        } else {
            const int n = min(16, w - x0);
            for (int p = 0; p < n; ++p)
                dst[p] = (uint8_t)bgr_luma(src[3 * p], src[3 * p + 1], src[3 * p + 2]);
        }
    }
}

// ------------------------------------------------------------------------------------------
// Frame table.  One persistent launch serves up to kMaxFrames frames of one parameter class
// (crop, neighbours, window), each at its own stage: random init (RandomInitializationForFrame,
// AdaptMIPMFrame.cu:474-518), one red/black phase (CheckerboardUpdateFrame, :520-947) or the final
// disparity record.  The frames' MIs form one index space; a single-frame table is the plain
// per-frame path (the reference's launch-per-phase pattern, :949-995), a multi-frame table the
// batched and served forms.  Descriptors travel by value in the kernel arguments, so a captured
// hipGraph replays them without memory nodes.
// ------------------------------------------------------------------------------------------
enum : int { STAGE_IDLE = 0, STAGE_INIT = 1, STAGE_PHASE = 2, STAGE_FINAL = 3 };
constexpr int kMaxFrames = 16;

struct FrameDesc {
    const float* tiles; const float2* centers; const int* nbr; const int* cnt;
    const float4* plane_in; const float* cost_in; const uint32_t* sel_in;   // phase / final input
    float4* plane_out; float* cost_out; uint32_t* sel_out;                  // init / phase output
    float4* disp;                                                           // final output
    // ACMM variant: top-k metadata after every update (and, in the final stage, cost/sel of the
    // final planes through cost_out/sel_out); warm levels carry the seeds of the coarser level
    const float4* plane_init; const float* cost_init; const uint32_t* sel_init;
    int acmm, warm;
    int V, mi_begin, stage, iter, phase;
    uint32_t seed, uid;
    // cost-vector records of the solve (wave engine; null = none), 5 + K floats per pixel:
    // [valid, plane.x, plane.y, plane.z, plane.w, cost_0 .. cost_{K-1}]
    float* cache;
};

struct StageArgs {
    FrameDesc f[kMaxFrames];
    int nf, total;
    unsigned long long* counters;  // counting variant: {samples, triples}
    unsigned* tickets;  // wave engine: 8 zeroed per-XCD MI counters (null: fixed MI stride)
    DevParams P;
};

__device__ __forceinline__ void active_pixel(int W, int a, int phase, int& x, int& y) {
    if (W & 1) {
        const int i = 2 * a + phase;
        y = i / W;
        x = i - y * W;
    } else {
        const int half = W >> 1;
        y = a / half;
        x = 2 * (a - y * half) + ((y + phase) & 1);
    }
}
// ------------------------------------------------------------------------------------------
// Wave-per-pixel solver (default).  One wavefront owns one pixel at a time and evaluates a
// task list of (hypothesis, view) pairs — every view whose anchor stays inside the crop, for
// every hypothesis of a batch — in passes of 8 tasks x 8 window-sample lanes.  The pixel's
// valid window samples (reference intensity, bilateral weight) are compacted once into LDS,
// so lanes never idle on clipped window positions or on views whose anchor left the crop, and
// each task's six weighted moments are reduced across its 8 lanes with DPP.  Results follow
// the definitions of ComputeBilateralNCC_Frame (AdaptMIPMFrame.cu:73-179), with the moments
// summed in a fixed order.
// ------------------------------------------------------------------------------------------
namespace wv {

constexpr int kSlots = 8;      // tasks per pass
constexpr int kHMax = 16;      // hypotheses per batch
// k_stage_w workgroups: up to 12 wavefronts (3 per SIMD, the occupancy its registers allow)
constexpr int kStageMaxThreads = 768;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ int prefix_below(uint64_t bal) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
}

// (the permutations used read a lane of every lane's own row, so no lane falls back to the
// bound value; bound_ctrl lets the compiler fold the move into the add)
template <int CTRL>
__device__ __forceinline__ float dppf(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}

// Sum over the 8 lanes of a task slot (lanes 8q..8q+7); every lane gets the same bits.
// floor(v) as an integer in one instruction (v_cvt_flr_i32_f32; saturates, 0 for NaN)
__device__ __forceinline__ int floor_int(float v) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

__device__ __forceinline__ float red8(float v) {
    v += dppf<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dppf<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dppf<0x141>(v);  // row_half_mirror
    return v;
}

// Window grid: offsets -r, -r+inc, ..., <= r on each axis (the reference's i/j loops).
__device__ __forceinline__ int grid_side(const DevParams& P) { return 2 * P.r / P.inc + 1; }

struct WaveLds {
    float4* samp;      // compacted valid window samples {i, j, x = I - I_c, w}, zero-padded
    float4* hyp;       // kHMax hypotheses of the current batch
    float2* hab;       // kHMax disparity slopes (A, B) = (-nx/nz, -ny/nz) of the hypotheses
    float* cost;       // [kHMax][K] NCC cost of (hypothesis, view); 2 = not scored
    uint16_t* task;    // (h << 8) | k
    int* iaux;         // 2*K ints (draw weights, selected-view list)
    float* sblur;      // blur term: reference blur minus the centre blur per compacted sample
    float* spare;      // a word that stores of masked task slots go to
};

__host__ __device__ inline size_t a16(size_t x) { return (x + 15) & ~(size_t)15; }

__host__ __device__ inline size_t shared_core_bytes(const DevParams& P) {
    const int n = 2 * P.r + 1;
    return a16((size_t)(P.K + 1) * P.TS * P.tile_bytes) + a16((size_t)P.WW * 4) +
           a16((size_t)n * n * 4) + a16((size_t)P.K * 16) + a16((size_t)P.K * 4) + 16;
}

// Shared part: tiles, tables, and (blur term) the FP32 blur slots after them.
__host__ __device__ inline size_t shared_bytes(const DevParams& P) {
    return shared_core_bytes(P) + (P.blur ? a16((size_t)(P.K + 1) * P.TS * 4) : 0);
}

__host__ __device__ inline size_t wave_bytes(const DevParams& P) {
    const int m = 2 * P.r / P.inc + 1;
    const size_t ns = (size_t)((m * m + 7) & ~7);
    return a16(ns * 16) + a16((size_t)kHMax * 16) + a16((size_t)kHMax * 8) +
           a16((size_t)kHMax * P.K * 4) + a16((size_t)kHMax * P.K * 2) + a16((size_t)P.K * 8) +
           (P.blur ? a16(ns * 4) : 0) + 16;
}

__host__ __device__ inline size_t lds_total(const DevParams& P, int nwaves) {
    return shared_bytes(P) + (size_t)nwaves * wave_bytes(P);
}

// Shared part: the same regions and meaning as Lds (wts/slist unused).
__device__ __forceinline__ Lds carve_shared(unsigned char* base, const DevParams& P) {
    Lds L;
    const int n = 2 * P.r + 1;
    unsigned char* q = base;
    L.tiles = (float*)q;
    q += a16((size_t)(P.K + 1) * P.TS * P.tile_bytes);
    L.costp = (float*)q; q += a16((size_t)P.WW * 4);
    L.spat = (float*)q; q += a16((size_t)n * n * 4);
    L.step = (float2*)q; q += (size_t)P.K * 8;
    L.cdiff = (float2*)q; q += (size_t)P.K * 8;
    q = base + (a16((size_t)(P.K + 1) * P.TS * P.tile_bytes) + a16((size_t)P.WW * 4) +
                a16((size_t)n * n * 4) + a16((size_t)P.K * 16));
    L.nid = (int*)q;
    L.next = (unsigned*)(base + shared_core_bytes(P) - 16);
    L.cur = (int*)(base + shared_core_bytes(P) - 12);
    L.btiles = P.blur ? (float*)(base + shared_core_bytes(P)) : nullptr;
    return L;
}

__device__ __forceinline__ WaveLds carve_wave(unsigned char* base, const DevParams& P, int w) {
    const int m = 2 * P.r / P.inc + 1;
    const size_t ns = (size_t)((m * m + 7) & ~7);
    unsigned char* q = base + shared_bytes(P) + (size_t)w * wave_bytes(P);
    WaveLds S;
    S.samp = (float4*)q; q += a16(ns * 16);
    S.hyp = (float4*)q; q += a16((size_t)kHMax * 16);
    S.hab = (float2*)q; q += a16((size_t)kHMax * 8);
    S.cost = (float*)q; q += a16((size_t)kHMax * P.K * 4);
    S.task = (uint16_t*)q; q += a16((size_t)kHMax * P.K * 2);
    S.iaux = (int*)q; q += a16((size_t)P.K * 8);
    S.sblur = P.blur ? (float*)q : nullptr;
    if (P.blur) q += a16(ns * 4);
    S.spare = (float*)q;
    return S;
}

// Compact the pixel's valid window samples (reference point inside the crop) into S.samp, in
// the reference's window order (i outer, j inner), zero-padded to a multiple of 8.
template <typename TT>
__device__ int build_samples(const Lds& L, const DevParams& P, const WaveLds& S, int px, int py,
                             float rc, float rbc) {
    const int lane = lane_id();
    const int r = P.r, inc = P.inc, W = P.W, n = 2 * r + 1, m = grid_side(P);
    const float inv_sc = TileScale<TT>::inv_sc(P);
    // w / m as (w + 1/2) * (1/m) in FP32: exact, since (w + 1/2) / m lies at least 1/(2m) from
    // an integer and the two roundings move it by less than m * 2^-23 for w < m^2
    const float inv_m = 1.0f / (float)m;
    int total = 0;
    for (int b = 0; b < m * m; b += 64) {
        const int w = b + lane;
        bool valid = false;
        float4 rec = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float bv = 0.0f;
        if (w < m * m) {
            const int ii = (int)(((float)w + 0.5f) * inv_m), jj = w - ii * m;
            const int i = -r + ii * inc, j = -r + jj * inc;
            const int rx = px + i, ry = py + j;
            valid = (unsigned)rx < (unsigned)W && (unsigned)ry < (unsigned)W;
            if (valid) {
                const float x = ldt(tile_base<TT>(L) + ry * P.W1 + rx) - rc;
                const float wgt = __expf(L.spat[(j + r) * n + (i + r)] - fabsf(x) * inv_sc);
                rec = make_float4((float)i, (float)j, x, wgt);
                if (P.blur) bv = L.btiles[ry * P.W1 + rx] - rbc;
            }
        }
        const uint64_t bal = __ballot(valid);
        if (valid) {
            S.samp[total + prefix_below(bal)] = rec;
            if (P.blur) S.sblur[total + prefix_below(bal)] = bv;
        }
        total += __popcll(bal);
    }
    const int pad = ((total + 7) & ~7) - total;
    if (lane < pad) {
        S.samp[total + lane] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (P.blur) S.sblur[total + lane] = 0.0f;
    }
    return total;
}

// Append the tasks of hypothesis slot h (plane hv) over the views in vmask (k < c) whose anchor
// stays inside the crop (AdaptMIPMFrame.cu:104-108); clears the hypothesis' cost row to 2.
__device__ __forceinline__ int add_tasks(const Lds& L, const DevParams& P, const WaveLds& S,
                                         int h, float4 hv, uint32_t vmask, int c, int px, int py,
                                         int ntask) {
    const int lane = lane_id();
    const float Wf = (float)P.W;
    for (int k = lane; k < P.K; k += 64) S.cost[h * P.K + k] = 2.0f;
    const float A = -hv.x / hv.z, B = -hv.y / hv.z;
    const bool hyp_ok = fabsf(hv.z) >= P.cos_max_slant && isfinite(A) && isfinite(B);
    if (lane == 0) S.hab[h] = make_float2(A, B);
    bool act = false;
    if (lane < c && hyp_ok && ((vmask >> lane) & 1u) && L.nid[lane] >= 0) {
        const float2 st = L.step[lane];
        const float ax = (float)px + st.x * hv.w, ay = (float)py + st.y * hv.w;
        act = ax >= 0.0f && ax < Wf && ay >= 0.0f && ay < Wf;
    }
    const uint64_t bal = __ballot(act);
    if (act) S.task[ntask + prefix_below(bal)] = (uint16_t)((h << 8) | lane);
    return ntask + __popcll(bal);
}

// add_tasks for the rows h0 .. h0 + nh - 1 in `rows` (bit r = row h0 + r; planes already in
// S.hyp), built side by side: lane (g, k) of a round takes row h0 + r0 + g and view k.  Tasks
// are appended in (row, view) order, as consecutive add_tasks calls append them; every cost row
// of the range is cleared to 2 (unless `clear` is false: a second round over other views).
__device__ __forceinline__ int add_tasks_rows(const Lds& L, const DevParams& P,
                                              const WaveLds& S, int h0, int nh, uint32_t rows,
                                              uint32_t vmask, int c, int px, int py, int ntask,
                                              bool clear = true) {
    const int lane = lane_id();
    const float Wf = (float)P.W;
    if (clear)
        for (int i = lane; i < nh * P.K; i += 64) S.cost[h0 * P.K + i] = 2.0f;
    // slopes and admissibility of the rows, one row per lane (nh <= kHMax)
    bool rok = false;
    if (lane < nh) {
        const float4 hv = S.hyp[h0 + lane];
        const float A = -hv.x / hv.z, B = -hv.y / hv.z;
        S.hab[h0 + lane] = make_float2(A, B);
        rok = ((rows >> lane) & 1u) && fabsf(hv.z) >= P.cos_max_slant && isfinite(A) &&
              isfinite(B);
    }
    const uint64_t okrows = __ballot(rok);
    if (c <= 0) return ntask;
    const int per = 64 / c;
    const int g = lane / c, k = lane - g * c;
    const float2 st = L.step[k];
    const bool view_ok = g < per && ((vmask >> k) & 1u) && L.nid[k] >= 0;
    for (int r0 = 0; r0 < nh; r0 += per) {
        const int r = r0 + g;
        const int h = h0 + r;
        bool act = false;
        if (view_ok && r < nh && ((okrows >> r) & 1u)) {
            const float hw = S.hyp[h].w;
            const float ax = (float)px + st.x * hw, ay = (float)py + st.y * hw;
            act = ax >= 0.0f && ax < Wf && ay >= 0.0f && ay < Wf;
        }
        const uint64_t bal = __ballot(act);
        if (act) S.task[ntask + prefix_below(bal)] = (uint16_t)((h << 8) | k);
        ntask += __popcll(bal);
    }
    return ntask;
}

// Evaluate the task list: S.cost[h][k] for every task.  ns = valid window samples.  BLUR adds
// the blur term of ComputeMultiViewCostVector_Frame (MIStereo/AdaptMIPMFrame.cu:211-218; blur
// sums of ComputeBilateralNCC_Frame :140-144 over the samples that enter the moments):
// cost = min(2, ncc + lambda * (1 - exp(-0.007368 * (sum_ref_blur - sum_src_blur)^2))), with
// both sums taken about the centre pixel's blur rbc (the offsets cancel in the difference).
template <bool COUNT, typename TT, bool BLUR, int W1C>
__device__ void run_tasks(const Lds& L, const DevParams& P, const WaveLds& S, int ntask, int ns,
                          int px, int py, float rc, float rbc, WorkCount& wc) {
    const int lane = lane_id();
    const int q = lane >> 3, u = lane & 7;
    const int T = (ns + 7) >> 3;
    const float pxf = (float)px, pyf = (float)py;
    const float min_var = TileScale<TT>::min_var(P);
    const TT* tiles = tile_base<TT>(L);
    // tile row pitch: a compile-time constant in the kernel specialised for one crop (the
    // second row of a 2x2 footprint is then an immediate offset of the first), else P.W1
    const int W1 = W1C ? W1C : P.W1;
    const unsigned Wu = (unsigned)(W1 - 1);
    // the last cell whose 2x2 footprint lies in the padded tile: a rejected sample reads at its
    // cell clamped to this one, so every read stays inside the tile
    const unsigned maxcell = (unsigned)((W1 - 2) * W1 + W1 - 2);
    for (int t0 = 0; t0 < ntask; t0 += kSlots) {
        const int t = t0 + q;
        // slots past the last task set up task 0's map (no branch) and stay masked
        const int e = S.task[t < ntask ? t : 0];
        const int h = e >> 8, k = e & 255;
        const float hw = S.hyp[h].w;
        const float2 ab = S.hab[h];
        const float2 st = L.step[k];
        const int base = (int)__umul24((unsigned)(k + 1), (unsigned)P.TS);
        // source point of window offset (i, j): q = (px, py) + (i, j) + step * (w + A i + B j),
        // i.e. q = Q0 + Qi * i + Qj * j per axis
        const float Qx0 = fmaf(st.x, hw, pxf), Qxi = fmaf(st.x, ab.x, 1.0f), Qxj = st.x * ab.y;
        const float Qy0 = fmaf(st.y, hw, pyf), Qyi = st.y * ab.x, Qyj = fmaf(st.y, ab.y, 1.0f);
        // a map with a non-finite coefficient puts every sample outside the crop (cost 2, the
        // value its row was cleared to); excluding it keeps every q below finite
        const bool live = t < ntask && __builtin_isfinite(Qx0) && __builtin_isfinite(Qxi) &&
                          __builtin_isfinite(Qxj) && __builtin_isfinite(Qy0) &&
                          __builtin_isfinite(Qyi) && __builtin_isfinite(Qyj);
        const TT* tsrc = tiles + base;
        float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f, a4 = 0.0f, a5 = 0.0f;
        float b0 = 0.0f, b1 = 0.0f;
        // window samples u, u + 8, ... in order, one per lane and step, two steps at a time:
        // both records are read, both cells located and every tap read before either sample
        // is accumulated (the floor conversion is inline asm, which neither the unroller nor
        // the scheduler moves loads across, so the overlap is written out here)
        struct Cell {
            float4 rec;
            float fx, fy;
            unsigned cell;
            bool ok;
        };
        auto locate = [&](const float4& rec) {
            Cell m;
            m.rec = rec;
            const float qx = fmaf(Qxj, rec.y, fmaf(Qxi, rec.x, Qx0));
            const float qy = fmaf(Qyj, rec.y, fmaf(Qyi, rec.x, Qy0));
            // cell and in-cell offset: floor(q) and q - floor(q) (exact for q >= 0)
            const int ix = floor_int(qx), iy = floor_int(qy);
            m.fx = __builtin_amdgcn_fractf(qx);
            m.fy = __builtin_amdgcn_fractf(qy);
            // 0 <= q < W on both axes
            m.ok = live && (unsigned)ix < Wu && (unsigned)iy < Wu;
            m.cell = min(__umul24((unsigned)iy, (unsigned)W1) + (unsigned)ix, maxcell);
            return m;
        };
        // taps: the 2x2 footprint as (top-left, top-right, bottom-left, bottom-right), or for
        // pair tiles as (top t, top d, bottom t, bottom d)
        auto taps = [&](const Cell& m, float4& t4, float4& b4) {
            const TT* tp = tsrc + m.cell;
            if constexpr (IsPair<TT>::value) {
                const TT p0 = tp[0], p1 = tp[W1];
                t4 = make_float4(pair_t(p0), pair_d(p0), pair_t(p1), pair_d(p1));
            } else {
                t4 = make_float4(ldt(tp), ldt(tp + 1), ldt(tp + W1), ldt(tp + W1 + 1));
            }
            if (BLUR) {
                const float* bp = L.btiles + base + m.cell;
                b4 = make_float4(bp[0], bp[1], bp[W1], bp[W1 + 1]);
            }
        };
        auto accumulate = [&](const Cell& m, int s, const float4& t4, const float4& b4) {
            const float4 rec = m.rec;
            const float fx = m.fx, fy = m.fy;
            const bool ok = m.ok;
            float top, bot;
            if constexpr (IsPair<TT>::value) {
                top = fmaf(fx, t4.y, t4.x);
                bot = fmaf(fx, t4.w, t4.z);
            } else {
                top = t4.x + fx * (t4.y - t4.x);
                bot = t4.z + fx * (t4.w - t4.z);
            }
            const float y = (top + fy * (bot - top)) - rc;
            if (BLUR) {
                const float btop = b4.x + fx * (b4.y - b4.x);
                const float bbot = b4.z + fx * (b4.w - b4.z);
                const float yb = (btop + fy * (bbot - btop)) - rbc;
                const bool okb = ok && s < ns;  // the zero padding has weight 0 but not blur 0
                b0 += okb ? S.sblur[s] : 0.0f;
                b1 += okb ? yb : 0.0f;
            }
            // a rejected sample has weight 0; its interpolant is finite (its reads are clamped
            // into the tile), so every product it enters is a zero that changes no sum
            const float wm = ok ? rec.w : 0.0f;
            const float wmx = wm * rec.z;
            const float wy = wm * y;
            if (COUNT && ok && s < ns) ++wc.samples;
            a0 += wm;
            a1 += wmx;
            a2 = fmaf(wmx, rec.z, a2);
            a3 += wy;
            a4 = fmaf(wy, y, a4);
            a5 = fmaf(wy, rec.z, a5);
        };
        int it = 0;
        for (; it + 2 <= T; it += 2) {
            const int s0 = (it << 3) | u, s1 = s0 + 8;
            const float4 r0 = S.samp[s0], r1 = S.samp[s1];
            const Cell m0 = locate(r0), m1 = locate(r1);
            float4 t0, t1, c0, c1;
            taps(m0, t0, c0);
            taps(m1, t1, c1);
            accumulate(m0, s0, t0, c0);
            accumulate(m1, s1, t1, c1);
        }
        if (it < T) {
            const int s0 = (it << 3) | u;
            const Cell m0 = locate(S.samp[s0]);
            float4 t0, c0;
            taps(m0, t0, c0);
            accumulate(m0, s0, t0, c0);
        }
        a0 = red8(a0); a1 = red8(a1); a2 = red8(a2);
        a3 = red8(a3); a4 = red8(a4); a5 = red8(a5);
        if (BLUR) { b0 = red8(b0); b1 = red8(b1); }
        // epilogue once per pass: hardware reciprocal / reciprocal square root (1 ulp), the
        // precision class of the reference's fast-math build
        // (computed in every lane without a branch; lanes failing the tests keep cost 2)
        const float inv = __builtin_amdgcn_rcpf(a0);
        const float mx = a1 * inv, my = a3 * inv;
        const float vx = a2 * inv - mx * mx;
        const float vy = a4 * inv - my * my;
        const float cov = a5 * inv - mx * my;
        const float ncc = fminf(2.0f, fmaxf(0.0f, 1.0f - cov * __builtin_amdgcn_rsqf(vx * vy)));
        float cst = live && a0 > 1e-6f && !(vx < min_var || vy < min_var) ? ncc : 2.0f;
        if (BLUR) {
            const float db = b0 - b1;
            const float bw = expf(-0.007368f * db * db);
            cst = live ? fminf(2.0f, fmaxf(0.0f, cst + P.blur_lambda * (1.0f - bw))) : cst;
        }
        // every lane of a live slot stores the same cost, the other slots store to a spare word
        // (a store without a branch keeps the reductions' last steps folded into their adds)
        float* dst = live ? S.cost + __umul24((unsigned)h, (unsigned)P.K) + k : S.spare;
        *dst = cst;
        if (COUNT && live && u == 0) ++wc.triples;
    }
    wave_sync();
}

// run_tasks of the kernel variant (BLUR: the frame carries blur tiles).  The blur variant is
// its own kernel instantiation so that its extra registers do not lower the occupancy of the
// kernel without it.
template <bool COUNT, typename TT, bool BLUR, int W1C>
__device__ __forceinline__ void run_batch(const Lds& L, const DevParams& P, const WaveLds& S,
                                          int ntask, int ns, int px, int py, float rc, float rbc,
                                          WorkCount& wc) {
    run_tasks<COUNT, TT, BLUR, W1C>(L, P, S, ntask, ns, px, py, rc, rbc, wc);
}

// Value of v in lane l; l is the same in every lane.
__device__ __forceinline__ int lane_int(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float lane_value(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// The selected views of an update, held in registers: lane j < n holds the view and draw weight
// of selection slot j; the total weight wn (summed in selection order) and RN(1/wn) are the same
// for every row.
struct Sel {
    int k;
    float w;
    float wn, winv;
    int n;
};

__device__ __forceinline__ Sel load_sel(const DevParams& P, const WaveLds& S, int nsel) {
    Sel sl;
    const int lane = lane_id();
    sl.n = nsel;
    sl.k = lane < nsel ? S.iaux[P.K + lane] : 0;
    sl.w = lane < nsel ? (float)S.iaux[sl.k] : 0.0f;
    float wn = 0.0f;
    for (int j = 0; j < nsel; ++j) wn += lane_value(sl.w, j);
    sl.wn = wn;
    sl.winv = wn > 0.0f ? 1.0f / wn : 0.0f;
    return sl;
}

// tot / wn correctly rounded: Markstein's correction of tot * RN(1/wn).  Here wn is a whole
// number of draws (1..32) and tot is 0 or in [2^-60, 64] (weights times costs in [0, 2]); for
// that domain the result equals the IEEE quotient for every float tot (checked exhaustively).
__device__ __forceinline__ float div_wn(float tot, float wn, float winv) {
    const float q0 = tot * winv;
    const float r = fmaf(-q0, wn, tot);
    return fmaf(r, winv, q0);
}

// Weighted cost of hypothesis row h over the selected views (ComputeWeightedCost_Frame
// :284-317), summed in selection order.
__device__ __forceinline__ float wsum(const DevParams& P, const WaveLds& S, const Sel& sl, int h) {
    float tot = 0.0f;
    for (int j = 0; j < sl.n; ++j) {
        const int k = lane_int(sl.k, j);
        const float w = lane_value(sl.w, j);
        tot += w * S.cost[h * P.K + k];
    }
    return sl.wn <= 1e-6f ? 2.0f : div_wn(tot, sl.wn, sl.winv);
}

// wsum over the first `slots` selection slots only (the other views counted as cost 0): the
// same terms added in the same order, so its value never exceeds the row's wsum (each FMA of a
// non-negative term rounds to at least the running sum, and the quotient is correctly rounded).
__device__ __forceinline__ float wsum_prefix(const DevParams& P, const WaveLds& S, const Sel& sl,
                                             int h, int slots) {
    float tot = 0.0f;
    for (int j = 0; j < slots; ++j) {
        const int k = lane_int(sl.k, j);
        const float w = lane_value(sl.w, j);
        tot += w * S.cost[h * P.K + k];
    }
    return sl.wn <= 1e-6f ? 2.0f : div_wn(tot, sl.wn, sl.winv);
}

// wsum of row `lane` for every lane below nrows (lanes at or above it score row 0): the rows
// of a batch summed side by side, each in selection order as wsum sums it.
__device__ __forceinline__ float wsum_lanes(const DevParams& P, const WaveLds& S, const Sel& sl,
                                            int nrows) {
    const int lane = lane_id();
    return wsum(P, S, sl, lane < nrows ? lane : 0);
}


// Rows 0 .. nrows-1 in `rows` (planes in S.hyp) scored over the selected views tsel, for a
// comparison against a best cost that is at most `bound`.  With 4 or more selected views the
// first round scores the first half of the selection; a row whose weighted cost over that half
// (wsum_prefix) already reaches `bound` is settled: its full weighted cost is at least as large,
// so it cannot win, and its other views are not scored (their entries stay 2, which keeps its
// wsum at or above `bound`).  The other rows get the remaining views in a second round, so
// every row that can win is scored on all of them, exactly as in one round.
template <bool COUNT, typename TT, bool BLUR, int W1C>
__device__ __forceinline__ void run_rows_bounded(const Lds& L, const DevParams& P,
                                                 const WaveLds& S, const Sel& sl, int nrows,
                                                 uint32_t rows, uint32_t tsel, int c, int x,
                                                 int y, int ns, float rc, float rbc,
                                                 float bound, WorkCount& wc) {
    const int lane = lane_id();
    if (sl.n < 4) {
        const int nt = add_tasks_rows(L, P, S, 0, nrows, rows, tsel, c, x, y, 0);
        wave_sync();
        run_batch<COUNT, TT, BLUR, W1C>(L, P, S, nt, ns, x, y, rc, rbc, wc);
        return;
    }
    const int half = (sl.n + 1) >> 1;
    // the views of selection slots 0 .. half-1 (slots hold the selected views in increasing
    // order)
    uint32_t m1 = 0u, rem = tsel;
    for (int j = 0; j < half; ++j) {
        m1 |= rem & (0u - rem);
        rem &= rem - 1u;
    }
    int nt = add_tasks_rows(L, P, S, 0, nrows, rows, m1, c, x, y, 0);
    wave_sync();
    run_batch<COUNT, TT, BLUR, W1C>(L, P, S, nt, ns, x, y, rc, rbc, wc);
    const float part = wsum_prefix(P, S, sl, lane < nrows ? lane : 0, half);
    const uint32_t open =
        (uint32_t)__ballot(lane < nrows && ((rows >> lane) & 1u) && part < bound);
    nt = add_tasks_rows(L, P, S, 0, nrows, open, tsel & ~m1, c, x, y, 0, false);
    wave_sync();
    run_batch<COUNT, TT, BLUR, W1C>(L, P, S, nt, ns, x, y, rc, rbc, wc);
    if (COUNT) {
        // counting only: score the settled rows' other views as well, into the skipped counts,
        // and put their entries back to 2
        const uint32_t settled = (uint32_t)__ballot(lane < nrows && ((rows >> lane) & 1u)) & ~open;
        nt = add_tasks_rows(L, P, S, 0, nrows, settled, tsel & ~m1, c, x, y, 0, false);
        wave_sync();
        WorkCount sk{0u, 0u, 0u, 0u};
        run_batch<COUNT, TT, BLUR, W1C>(L, P, S, nt, ns, x, y, rc, rbc, sk);
        wc.samples_skipped += sk.samples;
        wc.triples_skipped += sk.triples;
        const uint32_t vrest = tsel & ~m1;
        for (int i = lane; i < nrows * P.K; i += 64) {
            const int r = i / P.K, k = i - r * P.K;
            if (((settled >> r) & 1u) && ((vrest >> k) & 1u)) S.cost[i] = 2.0f;
        }
        wave_sync();
    }
}

// Rounded point of the cross-view proposal through view k at disparity w
// (BuildCrossViewPlaneProposal :362-376): its pixel index in the neighbour crop, -1 when the
// view has no neighbour or the point leaves the crop.
__device__ __forceinline__ int xview_point(const Lds& L, const DevParams& P, int k, int x, int y,
                                           float w) {
    if (L.nid[k] < 0) return -1;
    const float2 sk = L.step[k];
    const float qfx = (float)x + sk.x * w, qfy = (float)y + sk.y * w;
    const float fqx = floorf(qfx + 0.5f), fqy = floorf(qfy + 0.5f);
    const float Wf = (float)P.W;
    if (!(fqx >= 0.0f && fqx < Wf && fqy >= 0.0f && fqy < Wf)) return -1;
    return (int)fqy * P.W + (int)fqx;
}

// The proposal itself (:377-397): the neighbour's plane at q re-expressed in the reference MI.
__device__ __forceinline__ float4 xview_plane(const Lds& L, const DevParams& P,
                                              const float4* __restrict__ plane_in, int k, int x,
                                              int y, int q) {
    const int qx = q % P.W, qy = q / P.W;
    const float4 hn = plane_in[(size_t)L.nid[k] * P.WW + q];
    const float al = -hn.x / hn.z, be = -hn.y / hn.z;
    const float ga = hn.w - al * (float)qx - be * (float)qy;
    const float2 cdv = L.cdiff[k];
    float den = 1.0f + al * cdv.x + be * cdv.y;
    if (fabsf(den) < 1e-6f) den = den >= 0.0f ? 1e-6f : -1e-6f;
    const float aR = al / den, bR = be / den, cR = ga / den;
    float nx = -aR, ny = -bR, nz = 1.0f;
    normalize3(nx, ny, nz);
    return make_float4(nx, ny, nz, aR * (float)x + bR * (float)y + cR);
}

// Top-k support of cost row `row` (ComputeSelectedViewsTopK_Frame :227-282): mean of the k
// smallest costs below 2 (summed in ascending order), mask of the views at or under the k-th,
// and the first lowest-cost view.
__device__ __forceinline__ void topk_row(const DevParams& P, const WaveLds& S, int row, int c,
                                         float& mean, uint32_t& mask, int& bestk) {
    const float* cv = S.cost + row * P.K;
    int valid = 0;
    bestk = -1;
    float bv = 2.0f;
    for (int k = 0; k < c; ++k) {
        const float x = cv[k];
        if (x < 2.0f) {
            ++valid;
            if (bestk < 0 || x < bv) { bestk = k; bv = x; }
        }
    }
    mean = 2.0f;
    mask = 0u;
    if (valid <= 0) return;
    const int kk = min(valid, P.top_k);
    uint32_t taken = 0u;
    float sum = 0.0f, thr = 2.0f;
    for (int m = 0; m < kk; ++m) {
        int bi = -1;
        float b = 0.0f;
        for (int k = 0; k < c; ++k) {
            if ((taken >> k) & 1u) continue;
            if (bi < 0 || cv[k] < b) { bi = k; b = cv[k]; }
        }
        taken |= 1u << bi;
        sum += b;
        thr = b;
    }
    mean = sum / (float)kk;
    for (int k = 0; k < c; ++k)
        if (cv[k] <= thr) mask |= 1u << k;
}

// Normal of plane b rotated by three random angles in (-amp/2, amp/2), kept facing the camera
// (GeneratePerturbedNormalLF, CudaUtil.cu:105-142); draws angle_slot .. angle_slot + 2.
__device__ __forceinline__ float3 perturb_normal(DevStream& rs, int x, int y, float4 b,
                                                 float amp, uint32_t angle_slot) {
    float vx = (float)x, vy = (float)y, vz = 1.0f;
    normalize3(vx, vy, vz);
    // lane l < 3 takes angle l + 1: the three sines and cosines side by side (every lane active)
    // (both angle slot groups start a Philox block, so the three slots share it)
    static_assert(SLOT_REF_ANGLE % 4 == 0 && SLOT_WARM_ANGLE % 4 == 0, "angle slots in one block");
    const int l = lane_id() < 3 ? lane_id() : 0;
    const float a = (u01_dev(rs.raw_lane(angle_slot, (uint32_t)l)) - 0.5f) * amp;
    const float sa = sinf(a), ca = cosf(a);
    const float s1 = lane_value(sa, 0), s2 = lane_value(sa, 1), s3 = lane_value(sa, 2);
    const float c1 = lane_value(ca, 0), c2 = lane_value(ca, 1), c3 = lane_value(ca, 2);
    float px_ = (c2 * c3) * b.x + (c3 * s1 * s2 - c1 * s3) * b.y + (s1 * s3 + c1 * c3 * s2) * b.z;
    float py_ = (c2 * s3) * b.x + (c1 * c3 + s1 * s2 * s3) * b.y + (c1 * s2 * s3 - c3 * s1) * b.z;
    float pz_ = (-s2) * b.x + (c2 * s1) * b.y + (c1 * c2) * b.z;
    if (px_ * vx + py_ * vy + pz_ * vz >= 0.0f) { px_ = b.x; py_ = b.y; pz_ = b.z; }
    normalize3(px_, py_, pz_);
    return make_float3(px_, py_, pz_);
}

// Warm seed admissible for an ACMM level (WarmSeedIsValid_Frame, AdaptMIPMFrameACMM.cu:186-194).
__device__ __forceinline__ bool warm_valid(const DevParams& P, float4 h, float cost,
                                           uint32_t sel) {
    return h.w >= P.dmin && h.w <= P.dmax && cost < 1.0f && __popc(sel) >= 3;
}

// Cost-vector record of a pixel: the costs of its current plane over all views at this pixel,
// written by the update that chose the plane from fully scored rows and read by the pixel's next
// update (its current plane, row 0 of the first batch) and by the final record, which then skip
// scoring it again.  A record is used only when it is marked valid and holds the pixel's current
// plane bit for bit; the init stage writes every pixel's record, so within one solve a record
// stands for exactly the costs a fresh scoring of that plane gives.
constexpr uint32_t kRecValid = 1u;

// lane l < 5 loads header word l, lane k < K cost k (hdr 0 / cost 2 without records)
__device__ __forceinline__ void rec_load(const FrameDesc& F, const DevParams& P, size_t idx,
                                         float& hdr, float& cost) {
    hdr = 0.0f;
    cost = 2.0f;
    if (!F.cache) return;
    const float* r = F.cache + idx * (size_t)(5 + P.K);
    const int lane = lane_id();
    if (lane < 5) hdr = r[lane];
    if (lane < P.K) cost = r[5 + lane];
}

__device__ __forceinline__ bool rec_match(float hdr, float4 h) {
    return (uint32_t)lane_int(__float_as_int(hdr), 0) == kRecValid &&
           lane_int(__float_as_int(hdr), 1) == __float_as_int(h.x) &&
           lane_int(__float_as_int(hdr), 2) == __float_as_int(h.y) &&
           lane_int(__float_as_int(hdr), 3) == __float_as_int(h.z) &&
           lane_int(__float_as_int(hdr), 4) == __float_as_int(h.w);
}

// lane k < K stores cost k of plane h; an invalid record only clears the valid word
__device__ __forceinline__ void rec_store(const FrameDesc& F, const DevParams& P, size_t idx,
                                          float4 h, float ck, bool valid) {
    if (!F.cache) return;
    float* r = F.cache + idx * (size_t)(5 + P.K);
    const int lane = lane_id();
    if (lane == 0) r[0] = __uint_as_float(valid ? kRecValid : 0u);
    if (valid) {
        if (lane >= 1 && lane < 5) r[lane] = lane == 1 ? h.x : (lane == 2 ? h.y : (lane == 3 ? h.z : h.w));
        if (lane < P.K) r[5 + lane] = ck;
    }
}

// Random initialisation of one pixel (RandomInitializationForFrame :474-518); ACMM warm levels
// add the seed, a jittered seed and a seed with a perturbed normal
// (InitializeFrameFromRandomOrWarm, AdaptMIPMFrameACMM.cu:373-472).
template <bool COUNT, typename TT, bool BLUR, int W1C>
__device__ void init_pixel(const Lds& L, const DevParams& P, const FrameDesc& F,
                           const WaveLds& S, int v, int c, int pix, WorkCount& wc) {
    const int W = P.W, K = P.K;
    const int x = pix % W, y = pix / W;
    const size_t idx = (size_t)v * P.WW + pix;
    const float rc = ldt(tile_base<TT>(L) + y * P.W1 + x);
    const float rbc = P.blur ? L.btiles[y * P.W1 + x] : 0.0f;
    const int ns = build_samples<TT>(L, P, S, x, y, rc, rbc);
    DevStream st((uint32_t)idx, 0u, F.seed, F.uid);
    float4 h;
    h.w = u01_dev(st.raw(SLOT_INIT_DEPTH)) * (P.dmax - P.dmin) + P.dmin;
    random_normal(st, SLOT_INIT_MARSAGLIA, h.x, h.y, h.z);
    const uint32_t allv = c >= 32 ? 0xFFFFFFFFu : ((1u << c) - 1u);
    int nh = 1;
    float4 hs[4];
    hs[0] = h;
    if (F.acmm && F.warm) {
        const float4 warm = F.plane_init[idx];
        if (warm_valid(P, warm, F.cost_init[idx], F.sel_init[idx])) {
            float4 jit = warm;
            const float u = u01_dev(st.raw(SLOT_WARM_JITTER));
            jit.w = fminf(P.dmax, fmaxf(P.dmin, warm.w * (1.0f + (u - 0.5f) * 2.0f * 0.15f)));
            const float3 nr = perturb_normal(st, x, y, warm, 0.015f * 3.14159265358979323846f,
                                             SLOT_WARM_ANGLE);
            hs[1] = warm;
            hs[2] = jit;
            hs[3] = make_float4(nr.x, nr.y, nr.z, warm.w);
            nh = 4;
        }
    }
    if (lane_id() == 0)
        for (int i = 0; i < nh; ++i) S.hyp[i] = hs[i];
    wave_sync();
    const int nt = add_tasks_rows(L, P, S, 0, nh, 0xFFFFFFFFu, allv, c, x, y, 0);
    wave_sync();
    run_batch<COUNT, TT, BLUR, W1C>(L, P, S, nt, ns, x, y, rc, rbc, wc);
    // top-k mean and selection mask of each candidate; the lowest mean wins (first on ties)
    float mean;
    uint32_t mask;
    int bk;
    topk_row(P, S, 0, c, mean, mask, bk);
    int best = 0;
    for (int i = 1; i < nh; ++i) {
        float m2;
        uint32_t k2;
        int b2;
        topk_row(P, S, i, c, m2, k2, b2);
        if (m2 < mean) { mean = m2; mask = k2; best = i; }
    }
    const float4 hb = best == 0 ? hs[0] : (best == 1 ? hs[1] : (best == 2 ? hs[2] : hs[3]));
    rec_store(F, P, idx, hb, lane_id() < K ? S.cost[best * K + lane_id()] : 2.0f, true);
    if (lane_id() == 0) {
        F.plane_out[idx] = hb;
        F.cost_out[idx] = mean;
        F.sel_out[idx] = mask;
    }
}

// Lowest-cost view record of the final plane (ComputeWeightedCost_Frame's disp, :303-306).
template <bool COUNT, typename TT, bool BLUR, int W1C>
__device__ void final_pixel(const Lds& L, const DevParams& P, const FrameDesc& F,
                            const WaveLds& S, int v, int c, int pix, WorkCount& wc) {
    const int W = P.W;
    const int x = pix % W, y = pix / W;
    const size_t idx = (size_t)v * P.WW + pix;
    const float4 h = F.plane_in[idx];
    float rhdr, rcost;
    rec_load(F, P, idx, rhdr, rcost);
    if (F.cache && rec_match(rhdr, h)) {
        if (lane_id() < P.K) S.cost[lane_id()] = rcost;
        wave_sync();
    } else {
        const float rc = ldt(tile_base<TT>(L) + y * P.W1 + x);
        const float rbc = P.blur ? L.btiles[y * P.W1 + x] : 0.0f;
        const int ns = build_samples<TT>(L, P, S, x, y, rc, rbc);
        if (lane_id() == 0) S.hyp[0] = h;
        const uint32_t allv = c >= 32 ? 0xFFFFFFFFu : ((1u << c) - 1u);
        const int nt = add_tasks(L, P, S, 0, h, allv, c, x, y, 0);
        wave_sync();
        run_batch<COUNT, TT, BLUR, W1C>(L, P, S, nt, ns, x, y, rc, rbc, wc);
    }
    float mc = 2.0f;
    int am = -1;
    for (int k = 0; k < c; ++k) {
        const float cv = S.cost[k];
        if (cv < mc) { mc = cv; am = k; }
    }
    if (lane_id() == 0)
        F.disp[idx] = am >= 0 ? disp_record(L, P, am, h.w) : make_float4(0.f, 0.f, 0.f, 0.f);
    if (F.acmm) {  // PostRefineMetadataFrame (AdaptMIPMFrameACMM.cu:903-929)
        float mean;
        uint32_t mask;
        int bk;
        topk_row(P, S, 0, c, mean, mask, bk);
        if (lane_id() == 0) {
            F.cost_out[idx] = mean;
            F.sel_out[idx] = mask;
        }
    }
}

// Propagation candidates over lanes: lane 8s + j holds candidate j of slot s with the positions,
// conditions and order select_candidates visits them in (bits 0-3 dx + 8, 4-7 dy + 8, then the
// margins x >= xmin, x <= W-1-xoff, y >= ymin, y <= W-1-yoff in 4 bits each, bit 24 present).
// Each slot keeps its first lowest-cost candidate, as the sequential picks do; the border
// duplicates are applied after.
__constant__ uint32_t kCandLanes[64] = {
    0x1010078u, 0x1020168u, 0x1021068u, 0x1030257u, 0x1032059u, 0x0000000u, 0x0000000u, 0x0000000u,
    0x1030058u, 0x1040048u, 0x0000000u, 0x0000000u, 0x0000000u, 0x0000000u, 0x0000000u, 0x0000000u,
    0x1100098u, 0x12001a8u, 0x12010a8u, 0x13002b7u, 0x13020b9u, 0x0000000u, 0x0000000u, 0x0000000u,
    0x13000b8u, 0x14000c8u, 0x0000000u, 0x0000000u, 0x0000000u, 0x0000000u, 0x0000000u, 0x0000000u,
    0x1000187u, 0x1010286u, 0x1100286u, 0x1020375u, 0x1200395u, 0x0000000u, 0x0000000u, 0x0000000u,
    0x1000385u, 0x1000484u, 0x0000000u, 0x0000000u, 0x0000000u, 0x0000000u, 0x0000000u, 0x0000000u,
    0x1001089u, 0x101208au, 0x110208au, 0x102307bu, 0x120309bu, 0x0000000u, 0x0000000u, 0x0000000u,
    0x100308bu, 0x100408cu, 0x0000000u, 0x0000000u, 0x0000000u, 0x0000000u, 0x0000000u, 0x0000000u,
};

__device__ __forceinline__ void select_candidates_lanes(const float* __restrict__ costp, int W,
                                                        int x, int y, uint32_t ct, int pos[8],
                                                        int src[8]) {
    const int lane = lane_id();
    const int dx = (int)(ct & 15u) - 8, dy = (int)((ct >> 4) & 15u) - 8;
    const int m = (x - (int)((ct >> 8) & 15u)) | (W - 1 - x - (int)((ct >> 12) & 15u)) |
                  (y - (int)((ct >> 16) & 15u)) | (W - 1 - y - (int)((ct >> 20) & 15u));
    const bool ok = ((ct >> 24) & 1u) && m >= 0;
    const int p = (y + dy) * W + (x + dx);
    const float cv = ok ? costp[p] : __builtin_inff();
    float mn = fminf(cv, dppf<0xB1>(cv));
    mn = fminf(mn, dppf<0x4E>(mn));
    mn = fminf(mn, dppf<0x141>(mn));
    const uint64_t first = __ballot(ok && cv == mn);
    const uint64_t base = __ballot(ok && (lane & 7) == 0);
    const int center = y * W + x;
#pragma unroll
    for (int sl = 0; sl < 8; ++sl) {
        pos[sl] = center;
        src[sl] = -1;
        if ((base >> (8 * sl)) & 1u) {
            const int j = __builtin_ctzll((first >> (8 * sl)) & 0xFFull);
            pos[sl] = lane_int(p, 8 * sl + j);
            src[sl] = sl;
        }
    }
    const int H = W;
    if (y <= 2 && y > 0) { pos[1] = pos[0]; src[1] = 0; }
    if (y >= H - 3 && y < H - 1) { pos[3] = pos[2]; src[3] = 2; }
    if (x <= 2 && x > 0) { pos[5] = pos[4]; src[5] = 4; }
    if (x >= W - 3 && x < W - 1) { pos[7] = pos[6]; src[7] = 6; }
}

// One red/black update of one pixel (CheckerboardUpdateFrame :520-947).
template <bool COUNT, typename TT, bool BLUR, int W1C>
__device__ void phase_pixel(const Lds& L, const DevParams& P, const FrameDesc& F,
                            const WaveLds& S, int v, int c, int x, int y, uint32_t ctab,
                            WorkCount& wc) {
    const int lane = lane_id();
    const int W = P.W, WW = P.WW, K = P.K;
    const int phase = F.phase;
    const size_t base = (size_t)v * WW;
    const int center = y * W + x;
    const size_t idx = base + center;
    // view-selection priors from the 4 direct neighbours' previous selections (loaded first so
    // the loads overlap the window setup and the first batch)
    uint32_t dsel[4];
    bool dval[4];
    dval[0] = y > 0; dval[1] = y < W - 1; dval[2] = x > 0; dval[3] = x < W - 1;
    dsel[0] = dval[0] ? F.sel_in[idx - W] : 0u;
    dsel[1] = dval[1] ? F.sel_in[idx + W] : 0u;
    dsel[2] = dval[2] ? F.sel_in[idx - 1] : 0u;
    dsel[3] = dval[3] ? F.sel_in[idx + 1] : 0u;
    float rhdr, rcost;
    rec_load(F, P, idx, rhdr, rcost);
    const float rc = ldt(tile_base<TT>(L) + y * P.W1 + x);
    const float rbc = P.blur ? L.btiles[y * P.W1 + x] : 0.0f;
    const int ns = build_samples<TT>(L, P, S, x, y, rc, rbc);

    // batch 1: the current plane (row 0) and the propagation candidates (rows 1..)
    int pos[8], src[8];
    select_candidates_lanes(L.costp, W, x, y, ctab, pos, src);
    int row[8];   // cost row of the plane candidate slot cd reads, -1 if absent
    int nh = 1;
    const float4 hcur = F.plane_in[idx];
    int my_src = lane == 0 ? center : -1, my_row = 0;  // lane 1 + cd loads candidate cd
#pragma unroll
    for (int cd = 0; cd < 8; ++cd) {
        row[cd] = -1;
        if (src[cd] == cd) {
            row[cd] = nh;
            if (lane == cd + 1) { my_src = pos[cd]; my_row = nh; }
            ++nh;
        } else if (cd & 1) {
            if (src[cd] == cd - 1) row[cd] = row[cd - 1];  // border duplicate of the near slot
        }
    }
    if (my_src >= 0) S.hyp[my_row] = F.plane_in[base + my_src];
    wave_sync();
    // A row whose plane is bit-equal to an earlier row's (propagation copies planes unchanged,
    // so once the field settles neighbours often carry the same plane) has that row's costs at
    // this pixel: it is scored once and the candidate slots read the earlier row.
    int canon = lane;
    {
        // row g's plane is broadcast from lane g; every later row still its own canon that
        // matches it takes g (so the first equal row wins)
        const float4 a = S.hyp[lane < nh ? lane : 0];
        for (int g = 0; g + 1 < nh; ++g) {
            const int bx = lane_int(__float_as_int(a.x), g), by = lane_int(__float_as_int(a.y), g);
            const int bz = lane_int(__float_as_int(a.z), g), bw = lane_int(__float_as_int(a.w), g);
            if (lane > g && lane < nh && canon == lane && __float_as_int(a.x) == bx &&
                __float_as_int(a.y) == by && __float_as_int(a.z) == bz && __float_as_int(a.w) == bw)
                canon = g;
        }
    }
    const uint32_t own = (uint32_t)__ballot(lane < nh && canon == lane);
#pragma unroll
    for (int cd = 0; cd < 8; ++cd)
        if (row[cd] >= 0) row[cd] = lane_int(canon, row[cd]);
    const uint32_t allv = c >= 32 ? 0xFFFFFFFFu : ((1u << c) - 1u);
    // row 0 comes from the pixel's record when it holds the current plane
    const bool rec_hit = F.cache && rec_match(rhdr, hcur);
    const int nt = add_tasks_rows(L, P, S, 0, nh, rec_hit ? own & ~1u : own, allv, c, x, y, 0);
    if (rec_hit && lane < K) S.cost[lane] = rcost;
    wave_sync();
    run_batch<COUNT, TT, BLUR, W1C>(L, P, S, nt, ns, x, y, rc, rbc, wc);

    // sampling probabilities, lane k = view k
    const float thr = 0.8f * expf((float)(F.iter * F.iter) / (-90.0f));
    float pk = 0.0f;
    if (lane < c) {
        const int k = lane;
        float count = 0.0f, tmpw = 0.0f;
        int cf = 0;
#pragma unroll
        for (int cd = 0; cd < 8; ++cd) {
            if (row[cd] < 0) continue;
            const float cv = S.cost[row[cd] * K + k];
            if (cv < thr) { tmpw += expf(div_m018(cv * cv)); count += 1.0f; }
            if (cv > 1.2f) ++cf;
        }
        float pr = 0.0f;
        if (count > 2.0f && cf < 3) pr = tmpw / count;
        else if (cf < 3) pr = expf(thr * thr / (-0.32f));
        float prior = 0.0f;
#pragma unroll
        for (int ii = 0; ii < 4; ++ii)
            if (dval[ii]) prior += ((dsel[ii] >> k) & 1u) ? 0.9f : 0.1f;
        pk = pr * prior;
    }
    // cumulative distribution in view order (sequential, as the reference sums)
    float psum = 0.0f;
    for (int k = 0; k < c; ++k) psum += lane_value(pk, k);
    DevStream rs((uint32_t)idx, (uint32_t)(1 + 2 * F.iter + phase), F.seed, F.uid);
    uint32_t tsel = 0u;
    int nsel = 0;
    if (psum > 0.0f) {
        const float inv = 1.0f / psum;
        // lane s < 15 owns draw s: the first view whose cumulative probability exceeds it
        float draw = 2.0f;
        if (lane < (int)VIEW_DRAWS) {
            DevStream ls((uint32_t)idx, (uint32_t)(1 + 2 * F.iter + phase), F.seed, F.uid);
            draw = u01_dev(ls.raw(SLOT_VIEW_DRAW + lane)) - 1.1920928955078125e-07f;
        }
        int vs = -1;
        float cum = 0.0f;
        for (int k = 0; k < c; ++k) {
            cum += lane_value(pk, k) * inv;
            if (vs < 0 && cum > draw) vs = k;
        }
        // lane k (< c) counts the draws that chose view k
        int wk = 0;
        for (int d = 0; d < (int)VIEW_DRAWS; ++d) wk += lane_int(vs, d) == lane ? 1 : 0;
        const bool chosen = lane < c && wk > 0;
        const uint64_t bal = __ballot(chosen);
        tsel = (uint32_t)bal;
        nsel = __popcll(bal);
        if (chosen) S.iaux[K + prefix_below(bal)] = lane;
        if (lane < c) S.iaux[lane] = wk;
    } else {
        for (int k = lane; k < c; k += 64) S.iaux[k] = 0;
    }
    if (nsel == 0) {
        for (int k = lane; k < c; k += 64) {
            S.iaux[k] = 1;
            S.iaux[K + k] = k;
        }
        tsel = allv;
        nsel = c;
    }
    wave_sync();
    const Sel sl = load_sel(P, S, nsel);

    // current plane and propagation candidates under the selected views
    float4 best = hcur;
    float bcost;
    // all-view costs of the first batch's winner, stored as the pixel's record when neither the
    // cross-view nor the refinement candidates replace it (those are scored on the selected
    // views only)
    float keep = 2.0f;
    bool replaced = false;
    {
        const float wrow = wsum_lanes(P, S, sl, nh);
        bcost = lane_value(wrow, 0);
        int brow = 0;
#pragma unroll
        for (int cd = 0; cd < 8; ++cd) {
            if (row[cd] < 0) continue;
            const float cc = lane_value(wrow, row[cd]);
            if (cc < bcost) { bcost = cc; best = S.hyp[row[cd]]; brow = row[cd]; }
        }
        if (F.cache && lane < K) keep = S.cost[brow * K + lane];
    }

    // cross-view proposals through each selected view (BuildCrossViewPlaneProposal :350-398,
    // loop :907-936).  The reference builds proposal j at the rounded point of best.w with best
    // as updated by proposals < j.  Here the proposals of a batch are built at the batch's
    // starting best and scored together; the in-order replay keeps a proposal's score when its
    // rounded point is the same under the replayed best, and rebuilds and scores it alone (row
    // kHMax - 1) otherwise, so the outcome is the sequential loop's.
    constexpr int kSpec = kHMax - 1;
    for (int j0 = 0; j0 < nsel; j0 += kSpec) {
        const int jn = min(nsel - j0, kSpec);
        int qspec = -1, kj = 0;
        float4 pspec = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
        if (lane < jn) {
            kj = S.iaux[K + j0 + lane];
            qspec = xview_point(L, P, kj, x, y, best.w);
            if (qspec >= 0) pspec = xview_plane(L, P, F.plane_in, kj, x, y, qspec);
        }
        wave_sync();
        if (lane < jn && qspec >= 0) S.hyp[lane] = pspec;
        wave_sync();
        const uint32_t live = (uint32_t)__ballot(lane < jn && qspec >= 0);
        // bcost only falls during the replay, so the batch's starting value bounds it
        run_rows_bounded<COUNT, TT, BLUR, W1C>(L, P, S, sl, jn, live, tsel, c, x, y, ns, rc,
                                               rbc, bcost, wc);
        const float wrow = wsum_lanes(P, S, sl, jn);
        // lane j's rounded point under the replayed best (recomputed after each acceptance)
        int qcur = qspec;
        for (int j = 0; j < jn; ++j) {
            const int q = lane_int(qcur, j);
            if (q < 0) continue;
            float4 prop;
            float cc;
            if (q == lane_int(qspec, j)) {
                prop = S.hyp[j];
                cc = lane_value(wrow, j);
            } else {
                const int k = lane_int(kj, j);
                prop = xview_plane(L, P, F.plane_in, k, x, y, q);
                wave_sync();
                if (lane == 0) S.hyp[kSpec] = prop;
                const int n1 = add_tasks(L, P, S, kSpec, prop, tsel, c, x, y, 0);
                wave_sync();
                run_batch<COUNT, TT, BLUR, W1C>(L, P, S, n1, ns, x, y, rc, rbc, wc);
                cc = wsum(P, S, sl, kSpec);
            }
            if (cc < bcost) {
                bcost = cc;
                best = prop;
                replaced = true;
                if (lane > j && lane < jn) qcur = xview_point(L, P, kj, x, y, best.w);
            }
        }
    }

    // refinement: ACMP depth x normal candidates (AdaptMIPMUtil.cu:181-183)
    {
        const float pert = 0.02f;
        const float dnow = best.w;
        float drand;
        if (bcost < 0.1f)
            drand = (u01_dev(rs.raw(SLOT_REF_DEPTH_RAND)) - 0.5f) * 2.0f + dnow;
        else
            drand = u01_dev(rs.raw(SLOT_REF_DEPTH_RAND)) * (P.dmax - P.dmin) + P.dmin;
        float rnx, rny, rnz;
        random_normal(rs, SLOT_REF_MARSAGLIA, rnx, rny, rnz);
        const float3 pn = perturb_normal(rs, x, y, best, pert * 3.14159265358979323846f,
                                         SLOT_REF_ANGLE);
        const float px_ = pn.x, py_ = pn.y, pz_ = pn.z;
        const float lo = (1.0f - pert) * dnow, hi = (1.0f + pert) * dnow;
        const float dpert = u01_dev(rs.raw(SLOT_REF_DEPTH_PERT)) * (hi - lo) + lo;
        const float4 cur = best;
        wave_sync();
        if (lane < 5) {
            // (n_cur, d_rand), (n_rand, d_now), (n_rand, d_rand), (n_pert, d_now), (n_cur, d_pert)
            const int i = lane;
            float4 hc;
            hc.x = (i == 0 || i == 4) ? cur.x : (i == 3 ? px_ : rnx);
            hc.y = (i == 0 || i == 4) ? cur.y : (i == 3 ? py_ : rny);
            hc.z = (i == 0 || i == 4) ? cur.z : (i == 3 ? pz_ : rnz);
            hc.w = (i == 0 || i == 2) ? drand : (i == 4 ? dpert : dnow);
            S.hyp[i] = hc;
        }
        wave_sync();
        run_rows_bounded<COUNT, TT, BLUR, W1C>(L, P, S, sl, 5, 0x1Fu, tsel, c, x, y, ns, rc, rbc,
                                               bcost, wc);
        const float wrow = wsum_lanes(P, S, sl, 5);
        for (int i = 0; i < 5; ++i) {
            const float cc = lane_value(wrow, i);
            if (cc < bcost) { bcost = cc; best = S.hyp[i]; replaced = true; }
        }
    }
    if (F.acmm) {
        // DetailRestorer_FrameACMM (AdaptMIPMFrameACMM.cu:470-503) against the warm seed, then
        // the final plane's own top-k cost and selection (:881-898).  The reference's best_sel
        // is the sampled selection throughout the update.
        float4 seed = make_float4(0.f, 0.f, 1.f, 0.f);
        bool seed_ok = false;
        if (F.warm) {
            seed = F.plane_init[idx];
            seed_ok = warm_valid(P, seed, F.cost_init[idx], F.sel_init[idx]);
        }
        wave_sync();
        if (lane == 0) S.hyp[0] = best;
        int nst = add_tasks(L, P, S, 0, best, allv, c, x, y, 0);
        if (seed_ok) {
            if (lane == 0) S.hyp[1] = seed;
            nst = add_tasks(L, P, S, 1, seed, allv, c, x, y, nst);
        }
        wave_sync();
        run_batch<COUNT, TT, BLUR, W1C>(L, P, S, nst, ns, x, y, rc, rbc, wc);
        float fmean;
        uint32_t fmask;
        int fb;
        topk_row(P, S, 0, c, fmean, fmask, fb);
        if (seed_ok) {
            float scost;
            uint32_t ssel;
            int sb;
            topk_row(P, S, 1, c, scost, ssel, sb);
            const int seed_count = __popc(ssel), best_count = __popc(tsel);
            const float disp_gap = fabsf(best.w - seed.w);
            const float large_gap = 0.08f * fmaxf(P.base, 1e-3f);
            const bool clearly_better = (bcost + 0.10f < scost) || (best_count >= seed_count + 2);
            const bool ambiguous = bcost > scost - 0.03f;
            const bool maybe_carrier = disp_gap > large_gap && ambiguous && best_count <= seed_count;
            if ((!clearly_better && maybe_carrier) || (scost + 0.02f < bcost)) {
                best = seed;
                fmean = scost;
                fmask = ssel;
            }
        }
        bcost = fmean;
        tsel = fmask;
    }
    rec_store(F, P, idx, best, keep, !replaced && !F.acmm);
    if (lane == 0) {
        F.plane_out[idx] = best;
        F.cost_out[idx] = bcost;
        F.sel_out[idx] = tsel;
    }
    wave_sync();
}

// MI hand-out.  XCD x owns the MI range [x*total/8, (x+1)*total/8) (neighbouring MIs share
// tiles in the XCD's L2).  With tickets each workgroup takes the next MI of its XCD's range and,
// once that is exhausted, of the other XCDs' ranges in turn, so the workgroups finish together
// however the cost of their MIs varies; without them it walks its XCD's range with a fixed
// stride.  Which workgroup updates an MI does not change the result.  Called by wavefront 0
// (uniform control flow; lane 0 takes the tickets) and stored to L.cur; -1: none left.
__device__ __forceinline__ void next_mi(const Lds& L, unsigned* tickets, int total, int stride,
                                        int& turn, int& prev) {
    const int xcd = blockIdx.x & 7;
    int m = -1;
    if (!tickets) {
        const int hi = (int)(((long long)total * (xcd + 1)) >> 3);
        prev = prev < 0 ? (int)(((long long)total * xcd) >> 3) + (int)(blockIdx.x >> 3)
                        : prev + stride;
        m = prev < hi ? prev : -1;
    } else {
        for (; turn < 8; ++turn) {
            const int x = (xcd + turn) & 7;
            const int lo = (int)(((long long)total * x) >> 3);
            const int hi = (int)(((long long)total * (x + 1)) >> 3);
            unsigned r = 0u;
            if (lane_id() == 0) r = atomicAdd(&tickets[x], 1u);
            const int c = lo + (int)__builtin_amdgcn_readfirstlane(r);
            if (c < hi) {
                m = c;
                break;
            }
        }
    }
    if (lane_id() == 0) *L.cur = m;
}

// The next pixel index of the workgroup's MI for this wavefront (uniform).
__device__ __forceinline__ int next_pixel(const Lds& L) {
    unsigned r = 0u;
    if (lane_id() == 0) r = atomicAdd(L.next, 1u);
    return (int)__builtin_amdgcn_readfirstlane(r);
}

template <bool COUNT, typename TT, bool BLUR, int W1C>
__global__ void __launch_bounds__(kStageMaxThreads) __attribute__((amdgpu_waves_per_eu(3)))
k_stage_w(StageArgs a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const DevParams& P = a.P;
    const int NT = blockDim.x, t = threadIdx.x, WW = P.WW, W = P.W;
    // the wavefront index is uniform; saying so keeps the per-pixel index arithmetic and the
    // counter-based RNG on the scalar unit
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6), nwaves = NT >> 6;
    Lds L = carve_shared(lds_raw, P);
    const WaveLds S = carve_wave(lds_raw, P, wave);
    fill_spatial(L, P);
    // MIs come from next_mi; wavefront 0 fetches and the workgroup reads the result from L.cur
    unsigned* const tickets = a.tickets;
    const int total = a.total, stride = P.slots_per_xcd;
    int turn = 0, prev = -1;  // wavefront 0's hand-out state
    WorkCount wc{0u, 0u, 0u, 0u};
    int f = 0;  // frame of the current MI
    const uint32_t ctab = kCandLanes[t & 63];
    if (wave == 0) next_mi(L, tickets, total, stride, turn, prev);
    for (;;) {
        // every wavefront is done with the previous MI, and its successor is in L.cur
        __syncthreads();
        const int gm = __builtin_amdgcn_readfirstlane(*L.cur);
        if (gm < 0) break;
        while (f > 0 && gm < a.f[f].mi_begin) --f;  // MIs come in order within an XCD range
        while (f + 1 < a.nf && gm >= a.f[f + 1].mi_begin) ++f;
        const FrameDesc& F = a.f[f];
        const int stage = F.stage;
        if (stage == STAGE_IDLE) {  // uniform across the workgroup
            __syncthreads();  // every wavefront has read L.cur
            if (wave == 0) next_mi(L, tickets, total, stride, turn, prev);
            continue;
        }
        const int v = gm - F.mi_begin;
        const int c = stage_mi<TT>(L, P, v, F.tiles, F.centers, F.nbr, F.cnt,
                                    P.blur ? F.tiles + (size_t)F.V * WW : nullptr);
        if (stage == STAGE_PHASE)
            for (int i = t; i < WW; i += NT) L.costp[i] = F.cost_in[(size_t)v * WW + i];
        // pixels are handed out one at a time (wave w starts at pixel w), so the wavefronts
        // of the workgroup finish the MI together however their pixels' work varies; which
        // wavefront updates a pixel does not change its result
        if (t == 0) *L.next = (unsigned)nwaves;
        __syncthreads();
        if (wave == 0) next_mi(L, tickets, total, stride, turn, prev);  // L.cur was read
                                                                    // before the barrier
        if (stage == STAGE_PHASE) {
            const size_t base = (size_t)v * WW;
            const int phase = F.phase;
            // pass-through of the inactive colour (and of every pixel of an MI without views)
            for (int i = t; i < WW; i += NT) {
                const int x = i % W, y = i / W;
                if (((x + y) & 1) != phase || c <= 0) {
                    F.plane_out[base + i] = F.plane_in[base + i];
                    F.cost_out[base + i] = F.cost_in[base + i];
                    F.sel_out[base + i] = F.sel_in[base + i];
                }
            }
            if (c > 0) {
                const int nact = (W & 1) ? (phase == 0 ? (WW + 1) / 2 : WW / 2) : WW / 2;
                for (int aidx = wave; aidx < nact; aidx = next_pixel(L)) {
                    int x, y;
                    active_pixel(W, aidx, phase, x, y);
                    phase_pixel<COUNT, TT, BLUR, W1C>(L, P, F, S, v, c, x, y, ctab, wc);
                }
            }
        } else if (stage == STAGE_INIT) {
            for (int pix = wave; pix < WW; pix = next_pixel(L))
                init_pixel<COUNT, TT, BLUR, W1C>(L, P, F, S, v, c, pix, wc);
        } else {
            for (int pix = wave; pix < WW; pix = next_pixel(L))
                final_pixel<COUNT, TT, BLUR, W1C>(L, P, F, S, v, c, pix, wc);
        }
    }
    flush_counts<COUNT>(a.counters, wc);
}

}  // namespace wv

// ------------------------------------------------------------------------------------------
// Fixed-plane cost evaluation (hypotheses sorted by MI; one workgroup per MI)
// ------------------------------------------------------------------------------------------
struct EvalArgs {
    const float* tiles; const float2* centers; const int* nbr; const int* cnt;
    const int* hyp;   // n x 6 words
    const int* seg;   // V+1 CSR offsets into hyp
    int n;
    float* out;       // n x K
    unsigned long long* counters;
    DevParams P;
};

__global__ void k_segments(const int* __restrict__ hyp, int n, int V, int* __restrict__ seg) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v > V) return;
    int lo = 0, hi = n;  // first index with hyp.mi >= v
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (hyp[(size_t)mid * 6] < v) lo = mid + 1; else hi = mid;
    }
    seg[v] = lo;
}

// Wave-per-hypothesis cost evaluation: one wavefront scores one record over all
// of its MI's views with the task engine of wv::run_tasks.
template <bool COUNT, typename TT, bool BLUR, int W1C>
__global__ void __launch_bounds__(256) k_eval_w(EvalArgs a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const DevParams& P = a.P;
    const int NT = blockDim.x, W = P.W, WW = P.WW, K = P.K;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nwaves = NT >> 6;
    const int lane = threadIdx.x & 63;
    Lds L = wv::carve_shared(lds_raw, P);
    const wv::WaveLds S = wv::carve_wave(lds_raw, P, wave);
    fill_spatial(L, P);
    int first, end, stride;
    mi_range(P, first, end, stride);
    WorkCount wc{0u, 0u, 0u, 0u};
    for (int mi = first; mi < end; mi += stride) {
        const int h0 = a.seg[mi], h1 = a.seg[mi + 1];
        if (h0 >= h1) continue;  // uniform across the workgroup
        __syncthreads();
        const int c = stage_mi<TT>(L, P, mi, a.tiles, a.centers, a.nbr, a.cnt,
                                    P.blur ? a.tiles + (size_t)P.V * WW : nullptr);
        __syncthreads();
        for (int hi = h0 + wave; hi < h1; hi += nwaves) {
            const int* rec = a.hyp + (size_t)hi * 6;
            const int pix = rec[1];
            if (pix < 0 || pix >= WW || rec[0] != mi) {  // malformed record: no score
                for (int k = lane; k < K; k += 64) a.out[(size_t)hi * K + k] = 2.0f;
                continue;
            }
            const float4 h = make_float4(__int_as_float(rec[2]), __int_as_float(rec[3]),
                                         __int_as_float(rec[4]), __int_as_float(rec[5]));
            const int x = pix % W, y = pix / W;
            const float rc = ldt(tile_base<TT>(L) + y * P.W1 + x);
            const float rbc = P.blur ? L.btiles[y * P.W1 + x] : 0.0f;
            const int ns = wv::build_samples<TT>(L, P, S, x, y, rc, rbc);
            if (lane == 0) S.hyp[0] = h;
            const uint32_t allv = c >= 32 ? 0xFFFFFFFFu : ((1u << c) - 1u);
            const int nt = wv::add_tasks(L, P, S, 0, h, allv, c, x, y, 0);
            wv::wave_sync();
            wv::run_batch<COUNT, TT, BLUR, W1C>(L, P, S, nt, ns, x, y, rc, rbc, wc);
            for (int k = lane; k < K; k += 64) a.out[(size_t)hi * K + k] = S.cost[k];
            wv::wave_sync();
        }
    }
    flush_counts<COUNT>(a.counters, wc);
}

// ------------------------------------------------------------------------------------------
// Fused CheckModule pass (MIDisparityFilterCPU::FilterTile, MIDisparityFilter.cpp:423-693).
// FP32 with a fixed operation order and no contraction, so the reason mask is bit-identical to
// the CPU twin given identical inputs.  Snapshot form: all reads see the pre-filter state.
// Ordered form (FilterGlobal, MIDisparityFilter.cpp:973-985): an MI reads the filtered states of
// its neighbours with a smaller index and the pre-filter states of the others.
// ------------------------------------------------------------------------------------------
struct FilterArgs {
    const float* gray; const float2* centers; const int* nbr; const int* cnt;
    int V, W, K;
    float base;
    lfpm_filter_params f;
    const float4* plane_in; const float* cost_in; const float4* disp_in; const uint32_t* sel_in;
    float4* plane_out; float* cost_out; float4* disp_out; uint32_t* sel_out; uint8_t* reason;
    // ordered form: the MIs of the level this launch runs (mi_order[level_begin + block])
    const int* mi_order; int level_begin;
};

__device__ float f_bilin_disp(const FilterArgs& a, const float4* pl, int n, float x, float y) {
#pragma clang fp contract(off)
    const int W = a.W;
    if (x < 0 || y < 0 || x > W - 1 || y > W - 1) return __int_as_float(0x7fc00000);
    const int x0 = (int)floorf(x), y0 = (int)floorf(y);
    const int x1 = min(x0 + 1, W - 1), y1 = min(y0 + 1, W - 1);
    const float ax = x - (float)x0, ay = y - (float)y0;
    const size_t b = (size_t)n * W * W;
    float d[4];
    const int xi[4] = {x0, x0, x1, x1}, yi[4] = {y0, y1, y0, y1};
    for (int q = 0; q < 4; ++q) {
        float v = pl[b + yi[q] * W + xi[q]].w;
        if (!isfinite(v) || v <= a.f.min_valid_disp) return __int_as_float(0x7fc00000);
        d[q] = v;
    }
    return (1.0f - ax) * (1.0f - ay) * d[0] + (1.0f - ax) * ay * d[1] + ax * (1.0f - ay) * d[2] +
           ax * ay * d[3];
}

__device__ float f_bilin_cost(const FilterArgs& a, const float* co, int n, float x, float y) {
#pragma clang fp contract(off)
    const int W = a.W;
    if (x < 0 || y < 0 || x > W - 1 || y > W - 1) return __int_as_float(0x7fc00000);
    const int x0 = (int)floorf(x), y0 = (int)floorf(y);
    const int x1 = min(x0 + 1, W - 1), y1 = min(y0 + 1, W - 1);
    const float ax = x - (float)x0, ay = y - (float)y0;
    const size_t b = (size_t)n * W * W;
    const float c00 = co[b + y0 * W + x0], c01 = co[b + y1 * W + x0];
    const float c10 = co[b + y0 * W + x1], c11 = co[b + y1 * W + x1];
    if (!isfinite(c00) || !isfinite(c01) || !isfinite(c10) || !isfinite(c11))
        return __int_as_float(0x7fc00000);
    return (1.0f - ax) * (1.0f - ay) * c00 + (1.0f - ax) * ay * c01 + ax * (1.0f - ay) * c10 +
           ax * ay * c11;
}

__device__ bool f_bilin_gray(const FilterArgs& a, int v, float x, float y, float& out) {
#pragma clang fp contract(off)
    const int W = a.W;
    if (x < 0 || y < 0 || x > W - 1 || y > W - 1) return false;
    const int x0 = (int)floorf(x), y0 = (int)floorf(y);
    const int x1 = min(x0 + 1, W - 1), y1 = min(y0 + 1, W - 1);
    const float ax = x - (float)x0, ay = y - (float)y0;
    const float* g = a.gray + (size_t)v * W * W;
    const float v00 = g[y0 * W + x0], v01 = g[y1 * W + x0], v10 = g[y0 * W + x1],
                v11 = g[y1 * W + x1];
    out = (1.0f - ax) * (1.0f - ay) * v00 + (1.0f - ax) * ay * v01 + ax * (1.0f - ay) * v10 +
          ax * ay * v11;
    return true;
}

__device__ bool f_mean_std3(const FilterArgs& a, int v, float fx, float fy, float& mean,
                            float& sd) {
#pragma clang fp contract(off)
    const int W = a.W;
    const int cx = (int)roundf(fx), cy = (int)roundf(fy);
    if (cx < 1 || cy < 1 || cx >= W - 1 || cy >= W - 1) return false;
    const float* g = a.gray + (size_t)v * W * W;
    double s1 = 0.0, s2 = 0.0;
    for (int yy = cy - 1; yy <= cy + 1; ++yy)
        for (int xx = cx - 1; xx <= cx + 1; ++xx) {
            const double q = (double)g[yy * W + xx];
            s1 += q;
            s2 += q * q;
        }
    const double scale = 1.0 / 9.0;
    const double m = s1 * scale;
    const double var = fmax(s2 * scale - m * m, 0.0);
    mean = (float)m;
    sd = (float)sqrt(var);
    return isfinite(mean) && isfinite(sd) && sd > 1.0e-3f;
}

__device__ float f_photo(const FilterArgs& a, int v, int n, float px, float py, float qx,
                         float qy) {
#pragma clang fp contract(off)
    float Ir = 0.0f, In = 0.0f;
    if (!f_bilin_gray(a, v, px, py, Ir)) return __int_as_float(0x7fc00000);
    if (!f_bilin_gray(a, n, qx, qy, In)) return __int_as_float(0x7fc00000);
    float mr, sr, mn, sn;
    if (f_mean_std3(a, v, px, py, mr, sr) && f_mean_std3(a, n, qx, qy, mn, sn)) {
        const float ra = (Ir - mr) / sr;
        const float rb = (In - mn) / sn;
        return fabsf(ra - rb);
    }
    return fabsf(In - Ir);
}

// Views whose loads a pixel issues together before deciding on them (one memory round per
// group instead of one per view and function).
constexpr int kFilterViewGroup = 2;

template <bool ORDERED>
__global__ void __launch_bounds__(1024) k_filter(FilterArgs a) {
#pragma clang fp contract(off)
    __shared__ uint8_t keep1[LFPM_MAX_CROP * LFPM_MAX_CROP];
    const int t = threadIdx.x;
    int v;
    if (ORDERED) {
        // one MI of the current level; the levels before it were completed by earlier
        // launches
        v = a.mi_order[a.level_begin + (int)blockIdx.x];
        if (v < 0 || v >= a.V) return;  // uniform across the workgroup, before any barrier
    } else {
        // workgroups are dealt to the 8 XCDs in turn: each XCD takes a contiguous MI range, so
        // the neighbour states an MI reads are mostly read by the same XCD (one L2)
        const int per_xcd = (a.V + 7) / 8;
        v = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
        if (v >= a.V) return;  // uniform across the workgroup, before any barrier
    }
    const int W = a.W, WW = W * W;
    const lfpm_filter_params& fp = a.f;
    const int cntv = min(a.cnt[v], a.K);
    const size_t base = (size_t)v * WW;
    // the MI's neighbour ids and centres, shared by its pixels; the pixel's own state is loaded
    // before the barrier so both sets of loads are in flight together
    __shared__ int s_nb[LFPM_MAX_NEIGHBORS];
    __shared__ float2 s_cn[LFPM_MAX_NEIGHBORS];
    if (t < a.K) {  // the whole row, loaded beside the count (slots past it are never read)
        const int n = a.nbr[(size_t)v * a.K + t];
        s_nb[t] = n;
        s_cn[t] = (n >= 0 && n < a.V) ? a.centers[n] : make_float2(0.0f, 0.0f);
    }
    // the pixel's own state (kept for the outputs) and its disparity in LDS for the spike test
    __shared__ float s_d[LFPM_MAX_CROP * LFPM_MAX_CROP];
    float4 p0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), r0 = p0;
    float d = 0.0f, cost0 = 0.0f;
    uint32_t sel0 = 0u;
    // the pixel's own photometric terms, the same for every view
    float Ir = 0.0f, mr = 0.0f, sr = 0.0f;
    bool own3 = false;
    if (t < WW) {
        p0 = a.plane_in[base + t];
        r0 = a.disp_in[base + t];
        d = p0.w;
        s_d[t] = d;
        cost0 = a.cost_in[base + t];
        sel0 = a.sel_in[base + t];
        if (fp.enable_cycle_photo_filter) {
            const float px = (float)(t % W), py = (float)(t / W);
            f_bilin_gray(a, v, px, py, Ir);
            own3 = f_mean_std3(a, v, px, py, mr, sr);
        }
    }
    const float2 cv = a.centers[v];
    __syncthreads();
    uint8_t reason = LFPM_PR_KEEP;
    bool k1 = false;
    if (t < WW) {
        const int x = t % W, y = t / W;
        bool done = !isfinite(d) || d <= fp.min_valid_disp;
        if (!done && fp.enable_cost_filter) {
            const bool invalid = !isfinite(cost0) || cost0 < 0.0f;
            const bool high = isfinite(cost0) && cost0 > fp.max_cost;
            if (invalid || high) { reason = LFPM_PR_COST; done = true; }
        }
        // views in order: the first max_triplet set bits of the selection, or the first
        // max_triplet neighbours (bit b of m = neighbour slot b)
        uint32_t m = 0u;
        int nv = 0;
        if (!done) {
            if (fp.use_selected_views_only) {
                m = sel0;
                nv = min(__popc(m), fp.max_triplet);
            } else {
                nv = min(cntv, fp.max_triplet);
                m = nv >= 32 ? 0xFFFFFFFFu : ((1u << nv) - 1u);
            }
            if (nv == 0) { reason = LFPM_PR_NOSUPPORT; done = true; }
        }
        if (!done) {
            const float px = (float)x, py = (float)y;
            int support = 0, good = 0;
            constexpr int G = kFilterViewGroup;
            for (int g0 = 0; g0 < nv; g0 += G) {
                // phase 1: geometry of up to G views and every load they need, all in flight
                bool ok[G], ok3[G];
                int nn[G];
                float qxs[G], qys[G], dxs[G], dys[G], scs[G], axs[G], ays[G];
                float td[G][4], tc[G][4], tg[G][4], t9[G][9];
#pragma unroll
                for (int j = 0; j < G; ++j) {
                    ok[j] = false;
                    ok3[j] = false;
                    nn[j] = -1;
                    if (g0 + j < nv) {
                        const int b = __ffs(m) - 1;
                        m &= m - 1u;
                        const int n = b < cntv ? s_nb[b] : -1;
                        if (n >= 0) {
                            const float2 cn = s_cn[b];
                            const float Bx = cv.x - cn.x, By = cv.y - cn.y;
                            const float bl = sqrtf(Bx * Bx + By * By);
                            if (!(bl < 1.0e-6f)) {
                                const float dx = Bx / bl, dy = By / bl;
                                const float scale = bl / a.base;
                                const float qx = px + dx * d * scale;
                                const float qy = py + dy * d * scale;
                                if (!(qx < 0 || qy < 0 || qx > W - 1 || qy > W - 1)) {
                                    ok[j] = true;
                                    nn[j] = n;
                                    qxs[j] = qx; qys[j] = qy;
                                    dxs[j] = dx; dys[j] = dy; scs[j] = scale;
                                }
                            }
                        }
                    }
                    if (ok[j]) {
                        const int n = nn[j];
                        const int x0 = (int)floorf(qxs[j]), y0 = (int)floorf(qys[j]);
                        const int x1 = min(x0 + 1, W - 1), y1 = min(y0 + 1, W - 1);
                        axs[j] = qxs[j] - (float)x0;
                        ays[j] = qys[j] - (float)y0;
                        const int o[4] = {y0 * W + x0, y1 * W + x0, y0 * W + x1, y1 * W + x1};
                        const size_t nb0 = (size_t)n * WW;
                        const bool filtered = ORDERED && n < v;
                        const float4* pl = filtered ? a.plane_out : a.plane_in;
                        const float* co = filtered ? a.cost_out : a.cost_in;
#pragma unroll
                        for (int q = 0; q < 4; ++q) td[j][q] = pl[nb0 + o[q]].w;
                        if (fp.enable_cost_filter) {
#pragma unroll
                            for (int q = 0; q < 4; ++q) tc[j][q] = co[nb0 + o[q]];
                        }
                        if (fp.enable_cycle_photo_filter) {
                            const float* gn = a.gray + nb0;
#pragma unroll
                            for (int q = 0; q < 4; ++q) tg[j][q] = gn[o[q]];
                            const int cx = (int)roundf(qxs[j]), cy = (int)roundf(qys[j]);
                            ok3[j] = !(cx < 1 || cy < 1 || cx >= W - 1 || cy >= W - 1);
                            if (ok3[j]) {
#pragma unroll
                                for (int q = 0; q < 9; ++q)
                                    t9[j][q] = gn[(cy - 1 + q / 3) * W + cx - 1 + q % 3];
                            }
                        }
                    }
                }
                // phase 2: the decisions, view by view in order, with the operations and
                // their order of the per-view functions above
#pragma unroll
                for (int j = 0; j < G; ++j) {
                    if (!ok[j]) continue;
                    const float ax = axs[j], ay = ays[j];
                    bool dvalid = true;
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        dvalid = dvalid && isfinite(td[j][q]) && !(td[j][q] <= fp.min_valid_disp);
                    if (!dvalid) continue;
                    const float dn = (1.0f - ax) * (1.0f - ay) * td[j][0] +
                                     (1.0f - ax) * ay * td[j][1] + ax * (1.0f - ay) * td[j][2] +
                                     ax * ay * td[j][3];
                    if (!isfinite(dn)) continue;
                    if (fp.enable_cost_filter) {
                        float nc;
                        if (!isfinite(tc[j][0]) || !isfinite(tc[j][1]) || !isfinite(tc[j][2]) ||
                            !isfinite(tc[j][3]))
                            nc = __int_as_float(0x7fc00000);
                        else
                            nc = (1.0f - ax) * (1.0f - ay) * tc[j][0] +
                                 (1.0f - ax) * ay * tc[j][1] + ax * (1.0f - ay) * tc[j][2] +
                                 ax * ay * tc[j][3];
                        const bool inv = !isfinite(nc) || nc < 0.0f;
                        const bool hi = isfinite(nc) && nc > fp.max_cost;
                        if (inv || hi) continue;
                    }
                    ++support;
                    bool pg = true, pp = true;
                    if (fp.enable_cycle_geo_filter) {
                        const float phx = qxs[j] - dxs[j] * dn * scs[j];
                        const float phy = qys[j] - dys[j] * dn * scs[j];
                        const float ex = phx - px, ey = phy - py;
                        const float ge = sqrtf(ex * ex + ey * ey);
                        pg = isfinite(ge) && ge <= fp.max_geo_err_px;
                    }
                    if (fp.enable_cycle_photo_filter) {
                        const float In = (1.0f - ax) * (1.0f - ay) * tg[j][0] +
                                         (1.0f - ax) * ay * tg[j][1] +
                                         ax * (1.0f - ay) * tg[j][2] + ax * ay * tg[j][3];
                        float pe;
                        bool nb3 = false;
                        float mn = 0.0f, sn = 0.0f;
                        if (own3 && ok3[j]) {
                            double s1 = 0.0, s2 = 0.0;
#pragma unroll
                            for (int q = 0; q < 9; ++q) {
                                const double g = (double)t9[j][q];
                                s1 += g;
                                s2 += g * g;
                            }
                            const double scale = 1.0 / 9.0;
                            const double mm = s1 * scale;
                            const double var = fmax(s2 * scale - mm * mm, 0.0);
                            mn = (float)mm;
                            sn = (float)sqrt(var);
                            nb3 = isfinite(mn) && isfinite(sn) && sn > 1.0e-3f;
                        }
                        if (own3 && nb3) {
                            const float ra = (Ir - mr) / sr;
                            const float rb = (In - mn) / sn;
                            pe = fabsf(ra - rb);
                        } else {
                            pe = fabsf(In - Ir);
                        }
                        pp = isfinite(pe) && pe <= fp.max_photo_err_u;
                    }
                    if (pg && pp) ++good;
                }
            }
            if (support < fp.min_good_neighbors) reason = LFPM_PR_NOSUPPORT;
            else if (good < fp.min_good_neighbors) reason = LFPM_PR_CYCLE;
            else { k1 = true; reason = LFPM_PR_KEEP; }
        }
        keep1[t] = k1 ? 1 : 0;
    }
    __syncthreads();
    if (t < WW) {
        bool kf = k1;
        if (k1 && fp.enable_spike_filter) {
            const int x = t % W, y = t / W;
            // the kept 8-neighbours' disparities in fixed slots (+inf where none), sorted by an
            // 19-comparator network in registers: the sorted values, and so the median
            // vals[nvals / 2], are those of sorting the kept values alone
            float vals[8];
            int nvals = 0;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int dy = (q < 3) ? -1 : (q < 5 ? 0 : 1);
                const int dx = (q < 3) ? q - 1 : (q < 5 ? (q == 3 ? -1 : 1) : q - 6);
                const int xn = x + dx, yn = y + dy;
                vals[q] = __int_as_float(0x7f800000);
                if (xn < 0 || yn < 0 || xn >= W || yn >= W) continue;
                const int ni = yn * W + xn;
                if (!keep1[ni]) continue;
                const float dv = s_d[ni];
                if (!isfinite(dv) || dv <= fp.min_valid_disp) continue;
                vals[q] = dv;
                ++nvals;
            }
            if (nvals >= fp.spike_min_neighbors) {
                constexpr int net[19][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3}, {4, 6},
                                            {5, 7}, {1, 2}, {5, 6}, {0, 4}, {3, 7}, {1, 5}, {2, 6},
                                            {1, 4}, {3, 6}, {2, 4}, {3, 5}, {3, 4}};
#pragma unroll
                for (int c = 0; c < 19; ++c) {
                    const float lo = fminf(vals[net[c][0]], vals[net[c][1]]);
                    const float hi = fmaxf(vals[net[c][0]], vals[net[c][1]]);
                    vals[net[c][0]] = lo;
                    vals[net[c][1]] = hi;
                }
                float med = vals[0];
#pragma unroll
                for (int q = 1; q < 8; ++q)
                    if (q == nvals / 2) med = vals[q];
                if (isfinite(d) && fabsf(d - med) > fp.spike_abs_diff) {
                    kf = false;
                    reason = LFPM_PR_SPIKE;
                }
            }
        }
        const size_t g = base + t;
        float4 po = p0;
        float co = cost0;
        float4 dout = r0;
        uint32_t so = sel0;
        if (isfinite(d) && d > fp.min_valid_disp && !kf) {
            po = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
            co = 2.0f;
            dout = make_float4(-1.0f, -1.0f, 0.0f, 0.0f);
            if (fp.clear_selected_views_when_invalid) so = 0u;
        }
        a.plane_out[g] = po;
        a.cost_out[g] = co;
        a.disp_out[g] = dout;
        a.sel_out[g] = so;
        a.reason[g] = reason;
    }
}

// ------------------------------------------------------------------------------------------
// Focal-stack refocus
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_refocus(const float* __restrict__ tiles,
                                                 const float2* __restrict__ centers,
                                                 const int* __restrict__ l2m, int cols, int rows,
                                                 float pitch, int W, int out_w, int out_h, int S,
                                                 float vmin, float vmax, float* __restrict__ out) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int k = blockIdx.z;
    if (x >= out_w || y >= out_h) return;
    const float vk = S > 1 ? vmin + (vmax - vmin) * (float)k / (float)(S - 1) : vmin;
    const float halff = (float)(W - 1) * 0.5f;
    const float reach = vk * halff + 1.0f;
    // a / vk for the slice's fixed vk as a product with RN(1 / vk) and one fused correction:
    // the correctly rounded quotient (Markstein's theorem: y = RN(1/b), q0 within an ulp of
    // a/b, exact residual a - b q0 by FMA, then RN(q0 + r y)), the value of the division the
    // inclusion test below compares
    auto quot = [](float a, float b, float rb) {
        const float q0 = a * rb;
        return fmaf(fmaf(-q0, b, a), rb, q0);
    };
    const float rvk = 1.0f / vk;
    auto over_vk = [&](float a) { return quot(a, vk, rvk); };
    const float rs = pitch * 0.8660254037844386f;
    // lattice rows/columns whose centre can lie within `reach` (the snapped centre is within
    // 0.5 px of the lattice point, and reach carries 1 px of slack); the inclusion test below
    // decides, so any superset of these cells gives the same sums in the same order
    const float rrs = 1.0f / rs, rpitch = 1.0f / pitch;
    const int j0 = max(0, (int)floorf(quot(y - reach - 0.5f * pitch, rs, rrs)));
    const int j1 = min(rows - 1, (int)ceilf(quot(y + reach - 0.5f * pitch, rs, rrs)));
    const int WW = W * W;
    float acc = 0.0f;
    int nacc = 0;
    for (int j = j0; j <= j1; ++j) {
        const float off = 0.5f * pitch * (float)(1 + (j & 1));
        const int i0 = max(0, (int)floorf(quot(x - reach - off, pitch, rpitch)));
        const int i1 = min(cols - 1, (int)ceilf(quot(x + reach - off, pitch, rpitch)));
        for (int i = i0; i <= i1; ++i) {
            const int v = l2m[j * cols + i];
            if (v < 0) continue;
            const float2 c = centers[v];
            const float ux = over_vk((float)x - c.x) + halff;
            const float uy = over_vk((float)y - c.y) + halff;
            if (!(ux >= 0.0f && ux <= (float)(W - 1) && uy >= 0.0f && uy <= (float)(W - 1)))
                continue;
            const float fx0 = floorf(ux), fy0 = floorf(uy);
            const int x0 = (int)fx0, y0 = (int)fy0;
            const int x1 = min(x0 + 1, W - 1), y1 = min(y0 + 1, W - 1);
            const float fx = ux - fx0, fy = uy - fy0;
            const float* tl = tiles + (size_t)v * WW;
            const float a = tl[y0 * W + x0], b = tl[y0 * W + x1];
            const float cc = tl[y1 * W + x0], d = tl[y1 * W + x1];
            const float top = a + fx * (b - a), bot = cc + fx * (d - cc);
            acc += top + fy * (bot - top);
            ++nacc;
        }
    }
    out[((size_t)k * out_h + y) * out_w + x] = nacc ? acc / (float)nacc : 0.0f;
}

// Refocus, one workgroup per 16 x 16 output tile and all slices.  The candidate MIs of the tile
// (the lattice cells whose centre can reach it at the largest slice) are listed once, in
// lattice order.  Per slice the workgroup keeps the candidates whose projected crop can touch the
// tile and tabulates, for each of them, the projection of the tile's 16 columns and 16 rows
// (the inclusion test and the bilinear cell and weight per axis, from the same correctly rounded
// quotient as k_refocus); every pixel then sums its candidates in lattice order, so the output
// equals k_refocus's.  The host bounds the candidates of a tile (at most kRefocusCands) and
// sizes the LDS by that bound.
constexpr int kRefocusCands = 160;

struct RefocusAxis {
    int cell;    // floor(u), -1 when u is outside [0, W - 1]
    float frac;  // u - floor(u)
};

__global__ void __launch_bounds__(256) k_refocus_tile(const float* __restrict__ tiles,
                                                      const float2* __restrict__ centers,
                                                      const int* __restrict__ l2m, int cols,
                                                      int rows, float pitch, int W, int out_w,
                                                      int out_h, int S, float vmin, float vmax,
                                                      int cap, float* __restrict__ out) {
    // LDS sized by the host's bound `cap` on the candidates of a tile
    extern __shared__ __align__(16) unsigned char rf_lds[];
    RefocusAxis* s_tab = (RefocusAxis*)rf_lds;              // [cap][32]
    float2* s_c = (float2*)(s_tab + (size_t)cap * 32);      // [cap]
    int* s_id = (int*)(s_c + cap);                          // [cap]
    int* s_live = s_id + cap;                               // [cap]
    __shared__ int s_n, s_nl;
    // workgroups are dealt to the 8 XCDs in turn: give each XCD a contiguous run of tiles (row
    // major), so that the MI tiles its workgroups read stay in that XCD's L2
    const int tiles_x = (out_w + 15) / 16, ntiles = tiles_x * ((out_h + 15) / 16);
    const int per_xcd = (ntiles + 7) / 8;
    const int tile = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    if (tile >= ntiles) return;  // uniform across the workgroup, before any barrier
    const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
    const int X0 = (tile % tiles_x) * 16, Y0 = (tile / tiles_x) * 16;
    const int x = X0 + tx, y = Y0 + ty;
    const float halff = (float)(W - 1) * 0.5f;
    auto quot = [](float a, float b, float rb) {
        const float q0 = a * rb;
        return fmaf(fmaf(-q0, b, a), rb, q0);
    };
    auto slice_v = [&](int k) {
        return S > 1 ? vmin + (vmax - vmin) * (float)k / (float)(S - 1) : vmin;
    };
    const float rs = pitch * 0.8660254037844386f;
    const float rrs = 1.0f / rs, rpitch = 1.0f / pitch;
    // candidates: every lattice cell a pixel of the tile enumerates at any slice (the reach
    // grows with the slice's virtual depth), in lattice order; the row ranges first, then one
    // wavefront reads the cells' MIs and compacts them in order
    __shared__ int s_j0, s_nrow, s_i0[32], s_cnt[33];
    if (t == 0) {
        float vtop = vmin;
        for (int k = 0; k < S; ++k) vtop = fmaxf(vtop, slice_v(k));
        const float reach = vtop * halff + 1.0f;
        const int j0 = max(0, (int)floorf(quot((float)Y0 - reach - 0.5f * pitch, rs, rrs)));
        const int j1 = min(rows - 1,
                           (int)ceilf(quot((float)(Y0 + 15) + reach - 0.5f * pitch, rs, rrs)));
        const int nrow = min(max(j1 - j0 + 1, 0), 32);
        s_j0 = j0;
        s_nrow = nrow;
        s_cnt[0] = 0;
        for (int r = 0; r < nrow; ++r) {
            const int j = j0 + r;
            const float off = 0.5f * pitch * (float)(1 + (j & 1));
            const int i0 = max(0, (int)floorf(quot((float)X0 - reach - off, pitch, rpitch)));
            const int i1 = min(cols - 1,
                               (int)ceilf(quot((float)(X0 + 15) + reach - off, pitch, rpitch)));
            s_i0[r] = i0;
            s_cnt[r + 1] = s_cnt[r] + max(i1 - i0 + 1, 0);
        }
    }
    __syncthreads();
    if (t < 64) {
        const int nrow = s_nrow, ncell = s_cnt[nrow];
        int nn = 0;
        for (int q0 = 0; q0 < ncell; q0 += 64) {
            const int q = q0 + t;
            int v = -1;
            if (q < ncell) {
                int r = 0;
                while (s_cnt[r + 1] <= q) ++r;
                v = l2m[(s_j0 + r) * cols + s_i0[r] + (q - s_cnt[r])];
            }
            const bool keep = v >= 0;
            const uint64_t bal = __ballot(keep);
            const int at = nn + __popcll(bal & ((1ull << t) - 1ull));
            if (keep && at < cap) {
                s_id[at] = v;
                s_c[at] = centers[v];
            }
            nn += __popcll(bal);
        }
        if (t == 0) s_n = min(nn, cap);
    }
    __syncthreads();
    const int n = s_n;
    const int WW = W * W;
    for (int k = 0; k < S; ++k) {
        const float vk = slice_v(k);
        const float rvk = 1.0f / vk;
        // candidates whose projected crop [c - halff vk, c + halff vk] (with 1 px of slack) meets
        // the tile on both axes, kept in lattice order
        if (t < 64) {
            int nl = 0;
            const float ext = halff * vk + 1.0f;
            for (int c0 = 0; c0 < n; c0 += 64) {
                const int c = c0 + t;
                bool keep = false;
                if (c < n) {
                    const float2 cc = s_c[c];
                    keep = cc.x + ext >= (float)X0 && cc.x - ext <= (float)(X0 + 15) &&
                           cc.y + ext >= (float)Y0 && cc.y - ext <= (float)(Y0 + 15);
                }
                const uint64_t bal = __ballot(keep);
                if (keep) s_live[nl + __popcll(bal & ((1ull << t) - 1ull))] = c;
                nl += __popcll(bal);
            }
            if (t == 0) s_nl = nl;
        }
        __syncthreads();
        const int nl = s_nl;
        // projection tables of the kept candidates: entry r < 16 is column X0 + r, else row
        // Y0 + r - 16
        for (int e = t; e < nl * 32; e += 256) {
            const int li = e >> 5, r = e & 31;
            const float2 cc = s_c[s_live[li]];
            const float a = r < 16 ? (float)(X0 + r) - cc.x : (float)(Y0 + r - 16) - cc.y;
            const float u = quot(a, vk, rvk) + halff;
            RefocusAxis ax;
            if (u >= 0.0f && u <= (float)(W - 1)) {
                const float f0 = floorf(u);
                ax.cell = (int)f0;
                ax.frac = u - f0;
            } else {
                ax.cell = -1;
                ax.frac = 0.0f;
            }
            s_tab[li * 32 + r] = ax;
        }
        __syncthreads();
        float acc = 0.0f;
        int nacc = 0;
        for (int li = 0; li < nl; ++li) {
            const RefocusAxis ax = s_tab[li * 32 + tx], ay = s_tab[li * 32 + 16 + ty];
            if (ax.cell < 0 || ay.cell < 0) continue;
            const int x0 = ax.cell, y0 = ay.cell;
            const int x1 = min(x0 + 1, W - 1), y1 = min(y0 + 1, W - 1);
            const float* tl = tiles + (size_t)s_id[s_live[li]] * WW;
            const float a = tl[y0 * W + x0], b = tl[y0 * W + x1];
            const float c2 = tl[y1 * W + x0], d = tl[y1 * W + x1];
            const float top = fmaf(ax.frac, b - a, a), bot = fmaf(ax.frac, d - c2, c2);
            acc += fmaf(ay.frac, bot - top, top);
            ++nacc;
        }
        if (x < out_w && y < out_h)
            out[((size_t)k * out_h + y) * out_w + x] = nacc ? acc / (float)nacc : 0.0f;
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// V2R behavioural model fit, FP64: Gram-matrix partials, then a fixed-order sum and a
// column-scaled 3x3 solve with partial pivoting on one lane.
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_fit_partial(const float* __restrict__ aL,
                                                     const float* __restrict__ v, int n,
                                                     double* __restrict__ part) {
    __shared__ double red[9][256];
    double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double a = aL[i], vv = v[i];
        const double x0 = a * vv, x1 = vv, x2 = 1.0, yy = a;
        s[0] += x0 * x0; s[1] += x0 * x1; s[2] += x0 * x2;
        s[3] += x1 * x1; s[4] += x1 * x2; s[5] += x2 * x2;
        s[6] += x0 * yy; s[7] += x1 * yy; s[8] += x2 * yy;
    }
    for (int q = 0; q < 9; ++q) red[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int q = 0; q < 9; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 9; ++q) part[(size_t)blockIdx.x * 9 + q] = red[q][0];
}

__global__ void k_fit_solve(const double* __restrict__ part, int nblocks, double* __restrict__ c) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < nblocks; ++b)
        for (int q = 0; q < 9; ++q) s[q] += part[(size_t)b * 9 + q];
    double M[3][4] = {{s[0], s[1], s[2], s[6]}, {s[1], s[3], s[4], s[7]}, {s[2], s[4], s[5], s[8]}};
    double sc[3];
    for (int i = 0; i < 3; ++i) sc[i] = M[i][i] > 0 ? 1.0 / sqrt(M[i][i]) : 1.0;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) M[i][j] *= sc[i] * sc[j];
        M[i][3] *= sc[i];
    }
    for (int col = 0; col < 3; ++col) {
        int piv = col;
        for (int r = col + 1; r < 3; ++r)
            if (fabs(M[r][col]) > fabs(M[piv][col])) piv = r;
        if (piv != col)
            for (int q = 0; q < 4; ++q) { const double tmp = M[col][q]; M[col][q] = M[piv][q]; M[piv][q] = tmp; }
        for (int r = col + 1; r < 3; ++r) {
            const double f = M[r][col] / M[col][col];
            for (int q = col; q < 4; ++q) M[r][q] -= f * M[col][q];
        }
    }
    double z[3];
    for (int i = 2; i >= 0; --i) {
        double acc = M[i][3];
        for (int j = i + 1; j < 3; ++j) acc -= M[i][j] * z[j];
        z[i] = acc / M[i][i];
    }
    for (int i = 0; i < 3; ++i) c[i] = z[i] * sc[i];
}

__global__ void k_v2r_convert(const float* __restrict__ v, int n, double c0, double c1, double c2,
                              double* __restrict__ z) {
#pragma clang fp contract(off)
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double vv = v[i];
        const double zz = (vv * c1 + c2) / (1.0 - vv * c0);
        z[i] = (isnan(zz) || isinf(zz) || zz < 0.0) ? 0.0 : zz;
    }
}

// ------------------------------------------------------------------------------------------
// Launch plumbing
// ------------------------------------------------------------------------------------------
static int g_num_cus = 0;

static int device_cus() {
    if (g_num_cus > 0) return g_num_cus;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    g_num_cus = cus;
    return cus;
}

// LDS one workgroup may allocate on the current device (160 KiB on MI355X: a workgroup can hold
// the whole of a CU's LDS), read once per process.
static size_t g_lds_per_block = 0;
static size_t device_lds_per_block() {
    if (g_lds_per_block > 0) return g_lds_per_block;
    int dev = 0, b = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&b, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        b < 65536)
        b = 65536;
    g_lds_per_block = (size_t)b;
    return g_lds_per_block;
}

// Persistent grid over `mis` work items: a multiple of 8 workgroups (one share per XCD).
static int persistent_grid_n(int mis, int per_cu) {
    int g = device_cus() * per_cu;
    const int vv = ((mis + 7) / 8) * 8;
    g = std::min(g, std::max(8, vv));
    g = (g / 8) * 8;
    return std::max(8, g);
}
static int persistent_grid(const DevParams& P, int per_cu) { return persistent_grid_n(P.V, per_cu); }

static thread_local unsigned long long* g_counters = nullptr;

static const int kPhasePerCu = 8;    // k_stage_w workgroups per CU (persistent slots), default
static const int kMaxSlotsPerCu = 8; // upper bound of the autotuner's search

// Cost records of V MIs (5 + K floats per pixel).
static size_t rec_bytes(const DevParams& P, int V) {
    return (size_t)V * P.WW * (5 + P.K) * 4;
}
// The solver keeps its working state in LDS; its scratch is a 256-byte header (the per-XCD MI
// tickets) followed by the cost records of the P.V MIs.
static size_t stage_scratch_bytes(const DevParams& P, int grid) {
    (void)grid;
    return 256 + rec_bytes(P, P.V);
}
// Records of the MIs from `mi_offset` on in a scratch laid out as above (null: none).
static float* rec_region(const DevParams& P, void* scratch, int mi_offset) {
    if (!scratch) return nullptr;
    return (float*)((unsigned char*)scratch + 256) + (size_t)mi_offset * P.WW * (5 + P.K);
}
static size_t phase_scratch_bytes(const DevParams& P, int per_cu) {
    return stage_scratch_bytes(P, persistent_grid(P, per_cu));
}

static FrameDesc frame_desc(const float* tiles, const float* centers, const int* nbr,
                            const int* cnt, int V, uint32_t seed, uint32_t uid) {
    FrameDesc f;
    std::memset(&f, 0, sizeof(f));
    f.tiles = tiles; f.centers = (const float2*)centers; f.nbr = nbr; f.cnt = cnt;
    f.V = V; f.seed = seed; f.uid = uid;
    return f;
}

// Threads per workgroup of the wave engine: `waves` wavefronts (at most max_waves), or the
// largest of 12, 8, 6, 4, 2, 1 below it whose LDS fits what one workgroup may allocate (wide
// crops with many views and a wide window, or the blur term's second set of tiles).
static int wave_threads(const DevParams& P, int waves, int max_waves) {
    const int steps[6] = {12, 8, 6, 4, 2, 1};
    const int want = std::min(std::max(waves, 1), max_waves);
    for (int nw : steps)
        if (nw <= want && wv::lds_total(P, nw) <= device_lds_per_block()) return 64 * nw;
    return 64;
}

// Workgroups per CU that fit (registers, LDS) for a kernel at this block size and LDS.
template <typename KERN>
static int fit_per_cu(KERN kern, int threads, size_t lds) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, threads, lds) != hipSuccess || n < 1)
        return 1;
    return n;
}

// Wave-engine kernels of the counting / FP16-storage / blur variants (FP16 storage has no blur
// variant: check_pm_params refuses the pair).
using StageKernel = void (*)(StageArgs);
// kSpecialCrop: the crop of the kernel specialised for one tile pitch (FP32, no blur term)
static const int kSpecialCrop = 16;
template <typename TT>
static StageKernel stage_w_fmt(bool count, bool blur, int crop) {
    if (count) return blur ? wv::k_stage_w<true, TT, true, 0> : wv::k_stage_w<true, TT, false, 0>;
    if (!blur && crop == kSpecialCrop) return wv::k_stage_w<false, TT, false, kSpecialCrop + 1>;
    return blur ? wv::k_stage_w<false, TT, true, 0> : wv::k_stage_w<false, TT, false, 0>;
}
static StageKernel stage_w_kernel(bool count, int fmt, bool blur, int crop) {
    if (fmt == 3) return stage_w_fmt<PairH>(count, false, crop);
    return stage_w_fmt<float>(count, blur, crop);
}

__global__ void k_zero_tickets(unsigned* __restrict__ t) {
    if (threadIdx.x < 8) t[threadIdx.x] = 0u;
}

// Default wavefronts per k_stage_w workgroup: of 4, 8 and 12, the count that keeps the most
// wavefronts per CU (workgroups that fit by LDS and registers, at most kPhasePerCu, times their
// wavefronts), the smaller on ties.  Narrow crops fit three 4-wavefront workgroups per CU
// (12 wavefronts, the register limit); wide crops, whose tiles fill the LDS, get one or two
// larger workgroups instead.
static int auto_waves(const DevParams& P, StageKernel kern) {
    int best = 4, best_wpc = 0;
    for (int nw : {4, 8, 12}) {
        const int NT = wave_threads(P, nw, wv::kStageMaxThreads / 64);
        const int fit = std::min(fit_per_cu(kern, NT, wv::lds_total(P, NT / 64)), kPhasePerCu);
        const int wpc = fit * (NT / 64);
        if (wpc > best_wpc) {
            best_wpc = wpc;
            best = NT / 64;
        }
    }
    return best;
}

// Launch k_stage_w over the frames of `a` (mi_begin/total filled here).
static int launch_stage(StageArgs& a, int per_cu, int waves, void* scratch,
                        size_t scratch_bytes, hipStream_t s) {
    int total = 0;
    for (int i = 0; i < a.nf; ++i) {
        a.f[i].mi_begin = total;
        total += a.f[i].V;
    }
    a.total = total;
    if (total <= 0) return LFPM_OK;
    const StageKernel kern = stage_w_kernel(g_counters != nullptr, a.P.fmt,
                                            a.P.blur != 0, a.P.W);
    const int NT = wave_threads(a.P, waves > 0 ? waves : auto_waves(a.P, kern),
                                wv::kStageMaxThreads / 64);
    const size_t lds = wv::lds_total(a.P, NT / 64);
    if (lds > device_lds_per_block())
        return set_error(LFPM_E_UNSUPPORTED, "k_stage_w: %zu B of LDS per workgroup (crop %d, "
                         "%d views, radius %d) exceeds the device's %zu B", lds, a.P.W, a.P.K,
                         a.P.r, device_lds_per_block());
    const int fit = fit_per_cu(kern, NT, lds);
    const int grid = persistent_grid_n(total, std::max(1, std::min(per_cu, fit)));
    a.P.slots_per_xcd = grid / 8;
    a.counters = g_counters;
    // MI tickets in the scratch header, cleared by a kernel (not a memset node, so the
    // launch pair can sit in a captured graph); without a scratch the fixed MI stride
    a.tickets = nullptr;
    if (scratch && scratch_bytes >= 256) {
        a.tickets = (unsigned*)scratch;
        hipLaunchKernelGGL(k_zero_tickets, dim3(1), dim3(64), 0, s, a.tickets);
    }
    hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), lds, s, a);
    return launch_check("k_stage_w");
}
static size_t state_bytes(const DevParams& P) {
    return (size_t)P.V * P.WW * (16 + 4 + 4);
}

static size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Autotuned k_stage_w launch shape per parameter class (crop, neighbours, radius, tile format,
// blur term): persistent workgroups per CU and wavefronts per workgroup (the threads that share
// one MI's staged tiles).  Process-wide, guarded for callers on several host threads.
struct TuneKey {
    int crop, K, r, storage, blur;
    bool operator==(const TuneKey& o) const {
        return crop == o.crop && K == o.K && r == o.r && storage == o.storage && blur == o.blur;
    }
};
struct TuneVal {
    int per_cu, waves;
};
static std::vector<std::pair<TuneKey, TuneVal>> g_tuned;
static std::mutex g_tuned_mu;
static const int kDefaultWaves = 4;  // k_eval_w wavefronts per workgroup

static TuneKey tune_key(const lfpm_pm_params* p) {
    return TuneKey{p->crop, p->max_neighbors, p->radius, p->storage, p->blur_lambda > 0.0f};
}

static TuneVal stage_config(const lfpm_pm_params* p) {
    const TuneKey k = tune_key(p);
    std::lock_guard<std::mutex> g(g_tuned_mu);
    for (auto& e : g_tuned)
        if (e.first == k) return e.second;
    return TuneVal{kPhasePerCu, 0};
}

int stage_slots_per_cu(const lfpm_pm_params* p) { return stage_config(p).per_cu; }
int stage_waves(const lfpm_pm_params* p) { return stage_config(p).waves; }

int stage_waves_used(const lfpm_pm_params* p) {
    const DevParams P = make_dev_params(p, 1);
    const int w = stage_waves(p);
    const StageKernel kern = stage_w_kernel(false, P.fmt, P.blur != 0, P.W);
    return wave_threads(P, w > 0 ? w : auto_waves(P, kern), wv::kStageMaxThreads / 64) / 64;
}

void set_stage_config(const lfpm_pm_params* p, int per_cu, int waves) {
    const TuneKey k = tune_key(p);
    std::lock_guard<std::mutex> g(g_tuned_mu);
    for (auto& e : g_tuned)
        if (e.first == k) { e.second = TuneVal{per_cu, waves}; return; }
    g_tuned.push_back({k, TuneVal{per_cu, waves}});
}

void set_stage_slots_per_cu(const lfpm_pm_params* p, int per_cu) {
    set_stage_config(p, per_cu, stage_waves(p));
}

float* stage_records(const lfpm_pm_params* p, void* scratch, int mi_offset) {
    return rec_region(make_dev_params(p, std::max(mi_offset, 1)), scratch, mi_offset);
}

size_t stage_scratch_size(const lfpm_pm_params* p, int total_mis) {
    if (check_pm_params(p, std::max(total_mis, 1))) return 0;
    const DevParams P = make_dev_params(p, total_mis);
    return stage_scratch_bytes(P, persistent_grid_n(total_mis, stage_slots_per_cu(p)));
}

int stage_launch(const lfpm_pm_params* p, const StageFrame* frames, int n, void* scratch,
                 size_t scratch_bytes, void* stream) {
    int rc = check_pm_params(p, 1);
    if (rc) return rc;
    for (int i0 = 0; i0 < n; i0 += kMaxFrames) {
        StageArgs a;
        a.nf = std::min(kMaxFrames, n - i0);
        int total = 0;
        for (int i = 0; i < a.nf; ++i) {
            const StageFrame& s = frames[i0 + i];
            if (!s.tiles || !s.centers || !s.nbr || !s.cnt || s.V <= 0)
                return set_error(LFPM_E_INVALID, "stage: frame %d has no inputs", i0 + i);
            FrameDesc& f = a.f[i];
            f = frame_desc(s.tiles, s.centers, s.nbr, s.cnt, s.V, s.seed, s.uid);
            f.stage = s.stage; f.iter = s.iter; f.phase = s.phase;
            f.plane_in = (const float4*)s.plane_in; f.cost_in = s.cost_in; f.sel_in = s.sel_in;
            f.plane_out = (float4*)s.plane_out; f.cost_out = s.cost_out; f.sel_out = s.sel_out;
            f.disp = (float4*)s.disp;
            f.cache = s.rec;
            total += s.V;
        }
        a.P = make_dev_params(p, total);
        if ((rc = launch_stage(a, stage_slots_per_cu(p), stage_waves(p), scratch, scratch_bytes,
                               (hipStream_t)stream)))
            return rc;
    }
    return LFPM_OK;
}

} // namespace lfpm

using namespace lfpm;

extern "C" {

static int slice_impl(const uint8_t* raw, int raw_w, int raw_h, size_t raw_pitch, const int* org,
                      int V, int crop, float* out, float div, void* stream) {
    if (!raw || !org || !out || V <= 0 || crop < 1 || raw_w < crop || raw_h < crop ||
        raw_pitch < (size_t)raw_w)
        return set_error(LFPM_E_INVALID, "slice: bad arguments");
    const size_t total = (size_t)V * crop * crop;
    const int threads = 256;
    const int blocks = (int)std::min<size_t>((total + threads - 1) / threads, 65536);
    hipLaunchKernelGGL(k_slice, dim3(blocks), dim3(threads), 0, (hipStream_t)stream, raw,
                       raw_pitch, (const int2*)org, V, crop, div, out);
    return launch_check("slice");
}

int lfpm_slice_tiles(const uint8_t* raw_dev, int raw_w, int raw_h, size_t raw_pitch,
                     const int* origins_dev, int num_mis, int crop, float* tiles_dev,
                     void* stream) {
    return slice_impl(raw_dev, raw_w, raw_h, raw_pitch, origins_dev, num_mis, crop, tiles_dev,
                      255.0f, stream);
}

int lfpm_slice_gray(const uint8_t* raw_dev, int raw_w, int raw_h, size_t raw_pitch,
                    const int* origins_dev, int num_mis, int crop, float* gray_dev, void* stream) {
    return slice_impl(raw_dev, raw_w, raw_h, raw_pitch, origins_dev, num_mis, crop, gray_dev,
                      0.0f, stream);
}

int lfpm_slice_bgr_tiles(const uint8_t* bgr_dev, int raw_w, int raw_h, size_t bgr_pitch,
                         const int* origins_dev, int num_mis, int crop, uint8_t* bgr_tiles_dev,
                         void* stream) {
    if (!bgr_dev || !origins_dev || !bgr_tiles_dev || num_mis <= 0 || crop < 1 ||
        raw_w < crop || raw_h < crop || bgr_pitch < (size_t)raw_w * 3)
        return set_error(LFPM_E_INVALID, "slice_bgr: bad arguments");
    const size_t total = (size_t)num_mis * crop * crop;
    const int threads = 256;
    const int blocks = (int)std::min<size_t>((total + threads - 1) / threads, 65536);
    hipLaunchKernelGGL(k_slice_bgr, dim3(blocks), dim3(threads), 0, (hipStream_t)stream, bgr_dev,
                       bgr_pitch, (const int2*)origins_dev, num_mis, crop, bgr_tiles_dev);
    return launch_check("slice_bgr");
}

int lfpm_slice_blur_tiles(const uint8_t* raw_dev, int raw_w, int raw_h, size_t raw_pitch,
                          const int* origins_dev, int num_mis, int crop, float* blur_tiles_dev,
                          void* stream) {
    if (!raw_dev || !origins_dev || !blur_tiles_dev || num_mis <= 0 || crop < 1 ||
        raw_w < crop || raw_h < crop || raw_pitch < (size_t)raw_w)
        return set_error(LFPM_E_INVALID, "slice_blur: bad arguments");
    const size_t total = (size_t)num_mis * crop * crop;
    const int threads = 256;
    const int blocks = (int)std::min<size_t>((total + threads - 1) / threads, 65536);
    hipLaunchKernelGGL(k_slice_blur, dim3(blocks), dim3(threads), 0, (hipStream_t)stream,
                       raw_dev, raw_pitch, raw_w, raw_h, (const int2*)origins_dev, num_mis, crop,
                       blur_tiles_dev);
    return launch_check("slice_blur");
}

int lfpm_bgr_to_gray(const uint8_t* bgr_dev, int width, int height, size_t bgr_pitch,
                     uint8_t* gray_dev, size_t gray_pitch, void* stream) {
    if (!bgr_dev || !gray_dev || width < 1 || height < 1 || bgr_pitch < (size_t)width * 3 ||
        gray_pitch < (size_t)width)
        return set_error(LFPM_E_INVALID, "bgr_to_gray: bad arguments");
    const size_t total = (size_t)((width + 15) >> 4) * (size_t)height;
    const int threads = 256;
    const int blocks = (int)std::min<size_t>((total + threads - 1) / threads, 65536);
    hipLaunchKernelGGL(k_bgr_gray, dim3(blocks), dim3(threads), 0, (hipStream_t)stream, bgr_dev,
                       bgr_pitch, width, height, gray_dev, gray_pitch);
    return launch_check("bgr_to_gray");
}

size_t lfpm_patchmatch_workspace_size(int num_mis, const lfpm_pm_params* p) {
    if (check_pm_params(p, num_mis)) return 0;
    const DevParams P = make_dev_params(p, num_mis);
    // the largest occupancy the tuner may pick, so a workspace stays valid after tuning
    return align_up(state_bytes(P)) + align_up(phase_scratch_bytes(P, kMaxSlotsPerCu));
}

// Cost records of the step calls: kept in the workspace when it has room for them.
static float* step_records(const DevParams& P, void* ws, size_t ws_bytes) {
    if (!ws || ws_bytes < stage_scratch_bytes(P, 8)) return nullptr;
    return rec_region(P, ws, 0);
}

static int check_solver_ptrs(const float* tiles, const float* centers, const int* nbr,
                             const int* cnt) {
    if (!tiles || !centers || !nbr || !cnt) return set_error(LFPM_E_INVALID, "null input");
    return LFPM_OK;
}

int lfpm_patchmatch_init(const float* tiles_dev, const float* centers_dev,
                         const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                         int num_mis, const lfpm_pm_params* p, float* plane_dev, float* cost_dev,
                         uint32_t* sel_dev, void* workspace_dev, size_t workspace_bytes,
                         void* stream) {
    int rc = check_pm_params(p, num_mis);
    if (rc) return rc;
    if ((rc = check_solver_ptrs(tiles_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev)))
        return rc;
    if (!plane_dev || !cost_dev || !sel_dev) return set_error(LFPM_E_INVALID, "null output");
    StageArgs a;
    a.nf = 1;
    a.P = make_dev_params(p, num_mis);
    a.f[0] = frame_desc(tiles_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev, num_mis,
                        p->seed, p->frame_uid);
    a.f[0].stage = STAGE_INIT;
    a.f[0].plane_out = (float4*)plane_dev; a.f[0].cost_out = cost_dev; a.f[0].sel_out = sel_dev;
    a.f[0].cache = step_records(a.P, workspace_dev, workspace_bytes);
    return launch_stage(a, stage_slots_per_cu(p), stage_waves(p), workspace_dev, workspace_bytes,
                        (hipStream_t)stream);
}

int lfpm_patchmatch_phase(const float* tiles_dev, const float* centers_dev,
                          const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                          int num_mis, const lfpm_pm_params* p, int iter, int phase,
                          const float* plane_in_dev, const float* cost_in_dev,
                          const uint32_t* sel_in_dev, float* plane_out_dev, float* cost_out_dev,
                          uint32_t* sel_out_dev, void* workspace_dev, size_t workspace_bytes,
                          void* stream) {
    int rc = check_pm_params(p, num_mis);
    if (rc) return rc;
    if ((rc = check_solver_ptrs(tiles_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev)))
        return rc;
    if (!plane_in_dev || !cost_in_dev || !sel_in_dev || !plane_out_dev || !cost_out_dev ||
        !sel_out_dev)
        return set_error(LFPM_E_INVALID, "null state buffer");
    if (phase != 0 && phase != 1) return set_error(LFPM_E_INVALID, "phase must be 0 or 1");
    if (iter < 0) return set_error(LFPM_E_INVALID, "iter must be >= 0");
    StageArgs a;
    a.nf = 1;
    a.P = make_dev_params(p, num_mis);
    a.f[0] = frame_desc(tiles_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev, num_mis,
                        p->seed, p->frame_uid);
    FrameDesc& f = a.f[0];
    f.stage = STAGE_PHASE; f.iter = iter; f.phase = phase;
    f.plane_in = (const float4*)plane_in_dev; f.cost_in = cost_in_dev; f.sel_in = sel_in_dev;
    f.plane_out = (float4*)plane_out_dev; f.cost_out = cost_out_dev; f.sel_out = sel_out_dev;
    f.cache = step_records(a.P, workspace_dev, workspace_bytes);
    return launch_stage(a, stage_slots_per_cu(p), stage_waves(p), workspace_dev, workspace_bytes,
                        (hipStream_t)stream);
}

// The final stage of a solve on the solve's scratch (its cost records and MI tickets; null:
// score every pixel, fixed MI stride).
static int finalize_impl(const float* tiles_dev, const float* centers_dev,
                         const int* neighbor_ids_dev, const int* neighbor_counts_dev, int num_mis,
                         const lfpm_pm_params* p, const float* plane_dev, float* disp_dev,
                         void* scratch, size_t scratch_bytes, void* stream) {
    int rc = check_pm_params(p, num_mis);
    if (rc) return rc;
    if ((rc = check_solver_ptrs(tiles_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev)))
        return rc;
    if (!plane_dev || !disp_dev) return set_error(LFPM_E_INVALID, "null buffer");
    StageArgs a;
    a.nf = 1;
    a.P = make_dev_params(p, num_mis);
    a.f[0] = frame_desc(tiles_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev, num_mis,
                        p->seed, p->frame_uid);
    a.f[0].stage = STAGE_FINAL;
    a.f[0].plane_in = (const float4*)plane_dev; a.f[0].disp = (float4*)disp_dev;
    a.f[0].cache = step_records(a.P, scratch, scratch_bytes);
    return launch_stage(a, stage_slots_per_cu(p), stage_waves(p), scratch, scratch_bytes,
                        (hipStream_t)stream);
}

int lfpm_patchmatch_finalize(const float* tiles_dev, const float* centers_dev,
                             const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                             int num_mis, const lfpm_pm_params* p, const float* plane_dev,
                             float* disp_dev, void* stream) {
    return finalize_impl(tiles_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev, num_mis,
                         p, plane_dev, disp_dev, nullptr, 0, stream);
}

int lfpm_patchmatch_finalize_ws(const float* tiles_dev, const float* centers_dev,
                                const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                                int num_mis, const lfpm_pm_params* p, const float* plane_dev,
                                float* disp_dev, void* workspace_dev, size_t workspace_bytes,
                                void* stream) {
    return finalize_impl(tiles_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev, num_mis,
                         p, plane_dev, disp_dev, workspace_dev, workspace_bytes, stream);
}

int lfpm_patchmatch_frame(const float* tiles_dev, const float* centers_dev,
                          const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                          int num_mis, const lfpm_pm_params* p, float* plane_dev, float* cost_dev,
                          float* disp_dev, uint32_t* sel_dev, void* workspace_dev,
                          size_t workspace_bytes, int sync_each_phase, void* stream) {
    int rc = check_pm_params(p, num_mis);
    if (rc) return rc;
    const DevParams P = make_dev_params(p, num_mis);
    const size_t need = lfpm_patchmatch_workspace_size(num_mis, p);
    if (!workspace_dev || workspace_bytes < need)
        return set_error(LFPM_E_INVALID, "frame: workspace too small (%zu < %zu)", workspace_bytes,
                         need);
    const size_t npx = (size_t)P.V * P.WW;
    unsigned char* ws = (unsigned char*)workspace_dev;
    float* plane_b = (float*)ws;
    float* cost_b = (float*)(ws + npx * 16);
    uint32_t* sel_b = (uint32_t*)(ws + npx * 20);
    unsigned char* scratch = ws + align_up(state_bytes(P));
    const size_t scratch_bytes = workspace_bytes - align_up(state_bytes(P));
    hipStream_t s = (hipStream_t)stream;
    auto sync = [&]() -> int {
        if (!sync_each_phase) return LFPM_OK;
        HIP_TRY(hipStreamSynchronize(s));
        return LFPM_OK;
    };
    if ((rc = lfpm_patchmatch_init(tiles_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev,
                                   num_mis, p, plane_dev, cost_dev, sel_dev, scratch,
                                   scratch_bytes, stream)))
        return rc;
    if ((rc = sync())) return rc;
    for (int it = 0; it < p->iterations; ++it) {
        if ((rc = lfpm_patchmatch_phase(tiles_dev, centers_dev, neighbor_ids_dev,
                                        neighbor_counts_dev, num_mis, p, it, 0, plane_dev,
                                        cost_dev, sel_dev, plane_b, cost_b, sel_b, scratch,
                                        scratch_bytes, stream)))
            return rc;
        if ((rc = sync())) return rc;
        if ((rc = lfpm_patchmatch_phase(tiles_dev, centers_dev, neighbor_ids_dev,
                                        neighbor_counts_dev, num_mis, p, it, 1, plane_b, cost_b,
                                        sel_b, plane_dev, cost_dev, sel_dev, scratch,
                                        scratch_bytes, stream)))
            return rc;
        if ((rc = sync())) return rc;
    }
    if (disp_dev) {
        if ((rc = finalize_impl(tiles_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev,
                                num_mis, p, plane_dev, disp_dev, scratch, scratch_bytes,
                                stream)))
            return rc;
        if ((rc = sync())) return rc;
    }
    return LFPM_OK;
}

// ------------------------------------------------------------------------------------------
// ACMM multi-scale variant (AdaptMIPMFrameACMM.cpp:155-665, AdaptMIPMFrameACMM.cu:373-991)
// ------------------------------------------------------------------------------------------
struct AcmmLevel { int width; float scale; int radius, iterations, top_k; };

// Level schedule of BuildBlurAlignedPyramids / InitGPUParamsFromCPUParams
// (AdaptMIPMFrameACMM.cpp:155-200, 300-345): two levels only when the crop is at least 24 px.
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

static lfpm_pm_params acmm_level_params(const lfpm_pm_params* p, const AcmmLevel& l) {
    lfpm_pm_params q = *p;
    q.crop = l.width;
    q.base = p->base * l.scale;
    q.radius = l.radius;
    q.iterations = l.iterations;
    q.top_k = l.top_k;
    q.depth_min = p->depth_min * l.scale;
    q.depth_max = p->depth_max * l.scale;
    return q;
}

// Area-average resize of every tile W -> W0 (cv::resize INTER_AREA as OpenCV documents it:
// each output pixel is the mean of the source area it covers, partial pixels weighted by the
// covered fraction).  One thread per output pixel.
__global__ void k_area_resize(const float* __restrict__ src, int V, int W, int W0,
                              float* __restrict__ dst) {
    const long long n = (long long)V * W0 * W0;
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = (int)(i / (W0 * W0)), r = (int)(i - (long long)v * W0 * W0);
    const int dy = r / W0, dx = r - dy * W0;
    const float S = (float)W / (float)W0;
    const float ax = (float)dx * S, bx = fminf((float)(dx + 1) * S, (float)W);
    const float ay = (float)dy * S, by = fminf((float)(dy + 1) * S, (float)W);
    const int sx0 = (int)floorf(ax), sx1 = min(W - 1, (int)ceilf(bx) - 1);
    const int sy0 = (int)floorf(ay), sy1 = min(W - 1, (int)ceilf(by) - 1);
    const float* t = src + (size_t)v * W * W;
    float acc = 0.0f;
    for (int sy = sy0; sy <= sy1; ++sy) {
        const float wy = (fminf(by, (float)(sy + 1)) - fmaxf(ay, (float)sy)) / (by - ay);
        if (!(wy > 0.0f)) continue;
        float row = 0.0f;
        for (int sx = sx0; sx <= sx1; ++sx) {
            const float wx = (fminf(bx, (float)(sx + 1)) - fmaxf(ax, (float)sx)) / (bx - ax);
            if (wx > 0.0f) row += wx * t[sy * W + sx];
        }
        acc += wy * row;
    }
    dst[i] = acc;
}

__global__ void k_scale_centers(const float* __restrict__ c, int n, float s, float* __restrict__ o) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = c[i] * s;
}

// PrepareWarmStartFromPreviousLevel (AdaptMIPMFrameACMM.cpp:496-570): a fine pixel inherits the
// coarse plane under it (disparity scaled by the level ratio) when the coarse cost is below 1,
// at least 3 views were selected and the 3x3 disparity spread there is small; otherwise the
// marker plane (0, 0, 1, -1) with cost 2.
__global__ void k_warm_start(const float4* __restrict__ plane0, const float* __restrict__ cost0,
                             const uint32_t* __restrict__ sel0, int V, int W0, int W1,
                             float dscale, float thresh, float4* __restrict__ pinit,
                             float* __restrict__ cinit, uint32_t* __restrict__ sinit) {
    const long long n = (long long)V * W1 * W1;
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = (int)(i / (W1 * W1)), r = (int)(i - (long long)v * W1 * W1);
    const int y = r / W1, x = r - y * W1;
    const float sx = (float)W0 / (float)max(1, W1);
    const int py = min(W0 - 1, (int)floorf((float)y * sx));
    const int px = min(W0 - 1, (int)floorf((float)x * sx));
    const size_t pi = (size_t)v * W0 * W0 + (size_t)py * W0 + px;
    float mean = 0.0f, sq = 0.0f;
    for (int oy = -1; oy <= 1; ++oy) {
        const int yy = min(W0 - 1, max(0, py + oy));
        for (int ox = -1; ox <= 1; ++ox) {
            const int xx = min(W0 - 1, max(0, px + ox));
            const float d = plane0[(size_t)v * W0 * W0 + (size_t)yy * W0 + xx].w;
            mean += d;
            sq += d * d;
        }
    }
    mean /= 9.0f;
    const float var = fmaxf(0.0f, sq / 9.0f - mean * mean);
    const bool reliable = cost0[pi] < 1.0f && __popc(sel0[pi]) >= 3 && sqrtf(var) <= thresh;
    float4 h = make_float4(0.0f, 0.0f, 1.0f, -1.0f);
    float c = 2.0f;
    uint32_t m = 0u;
    if (reliable) {
        h = plane0[pi];
        h.w *= dscale;
        c = cost0[pi];
        m = sel0[pi];
    }
    pinit[i] = h;
    cinit[i] = c;
    sinit[i] = m;
}

// One ACMM level: init (random, or warm-seeded), iterations x 2 ACMM phases, then the
// post-refine metadata pass (disp record, top-k cost and selection) into plane/cost/sel/disp.
static int acmm_level_solve(const float* tiles, const float* centers, const int* nbr,
                            const int* cnt, int V, const lfpm_pm_params* q, float* plane,
                            float* cost, uint32_t* sel, float* disp, float* plane_b, float* cost_b,
                            uint32_t* sel_b, const float4* pinit, const float* cinit,
                            const uint32_t* sinit, hipStream_t s) {
    const int warm = pinit != nullptr;
    auto base_desc = [&](int stage) {
        StageArgs a;
        a.nf = 1;
        a.P = make_dev_params(q, V);
        a.f[0] = frame_desc(tiles, centers, nbr, cnt, V, q->seed, q->frame_uid);
        FrameDesc& f = a.f[0];
        f.stage = stage;
        f.acmm = 1;
        f.warm = warm;
        f.plane_init = pinit; f.cost_init = cinit; f.sel_init = sinit;
        return a;
    };
    int rc;
    {
        StageArgs a = base_desc(STAGE_INIT);
        a.f[0].plane_out = (float4*)plane; a.f[0].cost_out = cost; a.f[0].sel_out = sel;
        if ((rc = launch_stage(a, stage_slots_per_cu(q), stage_waves(q), nullptr, 0, s))) return rc;
    }
    for (int it = 0; it < q->iterations; ++it)
        for (int ph = 0; ph < 2; ++ph) {
            StageArgs a = base_desc(STAGE_PHASE);
            FrameDesc& f = a.f[0];
            f.iter = it; f.phase = ph;
            f.plane_in = (const float4*)(ph == 0 ? plane : plane_b);
            f.cost_in = ph == 0 ? cost : cost_b;
            f.sel_in = ph == 0 ? sel : sel_b;
            f.plane_out = (float4*)(ph == 0 ? plane_b : plane);
            f.cost_out = ph == 0 ? cost_b : cost;
            f.sel_out = ph == 0 ? sel_b : sel;
            if ((rc = launch_stage(a, stage_slots_per_cu(q), stage_waves(q), nullptr, 0, s))) return rc;
        }
    StageArgs a = base_desc(STAGE_FINAL);
    a.f[0].plane_in = (const float4*)plane;
    a.f[0].disp = (float4*)disp;
    a.f[0].cost_out = cost;
    a.f[0].sel_out = sel;
    return launch_stage(a, stage_slots_per_cu(q), stage_waves(q), nullptr, 0, s);
}

struct AcmmWs {
    size_t tiles0, centers0, plane0, cost0, sel0, disp0, plane0b, cost0b, sel0b;
    size_t pinit, cinit, sinit, plane_b, cost_b, sel_b, total;
};

// (blur: the coarse level's tiles carry its blur tiles after the intensities)
static AcmmWs acmm_ws(int V, int W, int W0, bool two, bool blur) {
    AcmmWs w{};
    const size_t n1 = (size_t)V * W * W, n0 = two ? (size_t)V * W0 * W0 : 0;
    size_t o = 0;
    auto take = [&](size_t b) { const size_t at = o; o += align_up(b); return at; };
    w.plane_b = take(n1 * 16); w.cost_b = take(n1 * 4); w.sel_b = take(n1 * 4);
    if (two) {
        w.tiles0 = take(n0 * 4 * (blur ? 2 : 1)); w.centers0 = take((size_t)V * 8);
        w.plane0 = take(n0 * 16); w.cost0 = take(n0 * 4); w.sel0 = take(n0 * 4);
        w.disp0 = take(n0 * 16);
        w.plane0b = take(n0 * 16); w.cost0b = take(n0 * 4); w.sel0b = take(n0 * 4);
        w.pinit = take(n1 * 16); w.cinit = take(n1 * 4); w.sinit = take(n1 * 4);
    }
    w.total = o;
    return w;
}

size_t lfpm_patchmatch_acmm_workspace_size(int num_mis, const lfpm_pm_params* p) {
    if (check_pm_params(p, num_mis)) return 0;
    AcmmLevel lv[2];
    const int L = acmm_levels(p->crop, lv);
    return acmm_ws(num_mis, p->crop, lv[0].width, L == 2, p->blur_lambda > 0.0f).total;
}

int lfpm_patchmatch_frame_acmm(const float* tiles_dev, const float* centers_dev,
                               const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                               int num_mis, const lfpm_pm_params* p, float* plane_dev,
                               float* cost_dev, float* disp_dev, uint32_t* sel_dev,
                               void* workspace_dev, size_t workspace_bytes, void* stream) {
    int rc = check_pm_params(p, num_mis);
    if (rc) return rc;
    if ((rc = check_solver_ptrs(tiles_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev)))
        return rc;
    if (!plane_dev || !cost_dev || !disp_dev || !sel_dev)
        return set_error(LFPM_E_INVALID, "acmm: null output");
    if (p->storage != 0)
        return set_error(LFPM_E_UNSUPPORTED, "acmm: level images are not 8-bit; storage must be 0");
    const bool blur = p->blur_lambda > 0.0f;
    AcmmLevel lv[2];
    const int L = acmm_levels(p->crop, lv);
    const AcmmWs w = acmm_ws(num_mis, p->crop, lv[0].width, L == 2, blur);
    if (!workspace_dev || workspace_bytes < w.total)
        return set_error(LFPM_E_INVALID, "acmm: workspace too small (%zu < %zu)", workspace_bytes,
                         w.total);
    unsigned char* ws = (unsigned char*)workspace_dev;
    hipStream_t s = (hipStream_t)stream;
    const int V = num_mis;
    float* plane_b = (float*)(ws + w.plane_b);
    float* cost_b = (float*)(ws + w.cost_b);
    uint32_t* sel_b = (uint32_t*)(ws + w.sel_b);
    const lfpm_pm_params q1 = acmm_level_params(p, lv[L - 1]);
    if (L == 1)
        return acmm_level_solve(tiles_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev, V,
                                &q1, plane_dev, cost_dev, sel_dev, disp_dev, plane_b, cost_b,
                                sel_b, nullptr, nullptr, nullptr, s);
    const int W = p->crop, W0 = lv[0].width;
    const lfpm_pm_params q0 = acmm_level_params(p, lv[0]);
    float* tiles0 = (float*)(ws + w.tiles0);
    float* centers0 = (float*)(ws + w.centers0);
    const long long n0 = (long long)V * W0 * W0, n1 = (long long)V * W * W;
    hipLaunchKernelGGL(k_area_resize, dim3((unsigned)((n0 + 255) / 256)), dim3(256), 0, s,
                       tiles_dev, V, W, W0, tiles0);
    // the level's blur images: the MIs' blur maps resized like the intensities
    // (BuildBlurAlignedPyramids, AdaptMIPMFrameACMM.cpp:338-350; the Gaussian alignment of
    // :295-302 is switched off in the reference)
    if (blur)
        hipLaunchKernelGGL(k_area_resize, dim3((unsigned)((n0 + 255) / 256)), dim3(256), 0, s,
                           tiles_dev + n1, V, W, W0, tiles0 + n0);
    hipLaunchKernelGGL(k_scale_centers, dim3((2 * V + 255) / 256), dim3(256), 0, s, centers_dev,
                       2 * V, lv[0].scale, centers0);
    if ((rc = launch_check("acmm level images"))) return rc;
    float* plane0 = (float*)(ws + w.plane0);
    float* cost0 = (float*)(ws + w.cost0);
    uint32_t* sel0 = (uint32_t*)(ws + w.sel0);
    if ((rc = acmm_level_solve(tiles0, centers0, neighbor_ids_dev, neighbor_counts_dev, V, &q0,
                               plane0, cost0, sel0, (float*)(ws + w.disp0),
                               (float*)(ws + w.plane0b), (float*)(ws + w.cost0b),
                               (uint32_t*)(ws + w.sel0b), nullptr, nullptr, nullptr, s)))
        return rc;
    float4* pinit = (float4*)(ws + w.pinit);
    float* cinit = (float*)(ws + w.cinit);
    uint32_t* sinit = (uint32_t*)(ws + w.sinit);
    const float dscale = lv[1].scale / std::max(lv[0].scale, 1e-6f);
    const float thresh = std::max(0.03f, 0.12f * (p->base * lv[0].scale));
    hipLaunchKernelGGL(k_warm_start, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, s,
                       (const float4*)plane0, cost0, sel0, V, W0, W, dscale, thresh, pinit, cinit,
                       sinit);
    if ((rc = launch_check("acmm warm start"))) return rc;
    return acmm_level_solve(tiles_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev, V, &q1,
                            plane_dev, cost_dev, sel_dev, disp_dev, plane_b, cost_b, sel_b, pinit,
                            cinit, sinit, s);
}

// Fixed-plane evaluation with the caller's segment table `seg` (num_mis + 1 ints, written
// here before the evaluation reads it).
static int eval_costs_with(const float* tiles_dev, const float* centers_dev,
                           const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                           int num_mis, const lfpm_pm_params* p, const void* hyp_dev, int n,
                           float* cost_out_dev, int* seg, hipStream_t s) {
    DevParams P = make_dev_params(p, num_mis);
    const int grid = persistent_grid(P, kPhasePerCu);
    P.slots_per_xcd = grid / 8;
    const int NT = wave_threads(P, kDefaultWaves, 4);
    const size_t lds = wv::lds_total(P, NT / 64);
    if (lds > device_lds_per_block())
        return set_error(LFPM_E_UNSUPPORTED, "eval: %zu B of LDS per workgroup exceeds the "
                         "device's %zu B", lds, device_lds_per_block());
    hipLaunchKernelGGL(k_segments, dim3((num_mis + 256) / 256), dim3(256), 0, s,
                       (const int*)hyp_dev, n, num_mis, seg);
    EvalArgs a;
    a.tiles = tiles_dev; a.centers = (const float2*)centers_dev; a.nbr = neighbor_ids_dev;
    a.cnt = neighbor_counts_dev; a.hyp = (const int*)hyp_dev; a.seg = seg; a.n = n;
    a.out = cost_out_dev; a.P = P; a.counters = g_counters;
    using EvalKernel = void (*)(EvalArgs);
    const bool blur = P.blur != 0;
    EvalKernel kern;
    if (P.fmt == 3)
        kern = g_counters ? k_eval_w<true, PairH, false, 0> : k_eval_w<false, PairH, false, 0>;
    else
        kern = g_counters ? (blur ? k_eval_w<true, float, true, 0> : k_eval_w<true, float, false, 0>)
                          : (blur ? k_eval_w<false, float, true, 0> : k_eval_w<false, float, false, 0>);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), lds, s, a);
    return launch_check("eval");
}

static int eval_costs_check(const float* tiles_dev, const float* centers_dev,
                            const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                            int num_mis, const lfpm_pm_params* p, const void* hyp_dev, int n,
                            float* cost_out_dev) {
    int rc = check_pm_params(p, num_mis);
    if (rc) return rc;
    if ((rc = check_solver_ptrs(tiles_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev)))
        return rc;
    if (n < 0 || (n > 0 && (!hyp_dev || !cost_out_dev)))
        return set_error(LFPM_E_INVALID, "eval: bad hypothesis buffer");
    return LFPM_OK;
}

int lfpm_eval_costs(const float* tiles_dev, const float* centers_dev,
                    const int* neighbor_ids_dev, const int* neighbor_counts_dev, int num_mis,
                    const lfpm_pm_params* p, const void* hyp_dev, int n, float* cost_out_dev,
                    void* stream) {
    int rc = eval_costs_check(tiles_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev,
                              num_mis, p, hyp_dev, n, cost_out_dev);
    if (rc || n == 0) return rc;
    hipStream_t s = (hipStream_t)stream;
    int* seg = nullptr;
    HIP_TRY(hipMallocAsync((void**)&seg, sizeof(int) * (num_mis + 1), s));
    rc = eval_costs_with(tiles_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev, num_mis,
                         p, hyp_dev, n, cost_out_dev, seg, s);
    const hipError_t fe = hipFreeAsync(seg, s);
    if (rc) return rc;
    if (fe != hipSuccess)
        return set_error(LFPM_E_HIP, "eval: hipFreeAsync: %s", hipGetErrorString(fe));
    return LFPM_OK;
}

// Allocation-free evaluation.  `scratch` holds the segment table (LFPM_NOALLOC_MAX_MIS + 1
// ints, 1 MiB); each device holds its own copy until the process ends.  Host code never passes
// the array's name to a kernel: it uses the device address taken once per device with
// hipGetSymbolAddress.  k_segments rewrites entries 0..num_mis on every call before the
// evaluation reads them, so nothing has to be cleared.  Calls that overlap on one device share
// the table and race.
static __device__ __align__(256) int scratch[LFPM_NOALLOC_MAX_MIS + 1];
static constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static int* g_scratch_ptr[kMaxScratchDevices] = {};

static int scratch_on_device(int** out) {
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices)
        return set_error(LFPM_E_UNSUPPORTED, "eval_noalloc: device %d beyond %d", dev,
                         kMaxScratchDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_error(LFPM_E_HIP, "eval_noalloc: no address for the scratch table");
        g_scratch_ptr[dev] = (int*)p;
    }
    *out = g_scratch_ptr[dev];
    return LFPM_OK;
}

int lfpm_eval_costs_noalloc(const float* tiles_dev, const float* centers_dev,
                            const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                            int num_mis, const lfpm_pm_params* p, const void* hyp_dev, int n,
                            float* cost_out_dev, void* stream) {
    int rc = eval_costs_check(tiles_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev,
                              num_mis, p, hyp_dev, n, cost_out_dev);
    if (rc) return rc;
    if (num_mis > LFPM_NOALLOC_MAX_MIS)
        return set_error(LFPM_E_INVALID, "eval_noalloc: %d MIs exceed the scratch table's %d",
                         num_mis, LFPM_NOALLOC_MAX_MIS);
    if (n == 0) return LFPM_OK;
    int* seg = nullptr;
    if ((rc = scratch_on_device(&seg))) return rc;
    return eval_costs_with(tiles_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev,
                           num_mis, p, hyp_dev, n, cost_out_dev, seg, (hipStream_t)stream);
}

static int filter_impl(bool ordered, const float* gray_dev, const float* centers_dev,
                       const int* neighbor_ids_dev, const int* neighbor_counts_dev, int num_mis,
                       int crop, int max_neighbors, float base, const lfpm_filter_params* f,
                       const float* plane_in_dev, const float* cost_in_dev,
                       const float* disp_in_dev, const uint32_t* sel_in_dev,
                       float* plane_out_dev, float* cost_out_dev, float* disp_out_dev,
                       uint32_t* sel_out_dev, uint8_t* reason_dev, const int* mi_order_dev,
                       const int* level_start, int num_levels, hipStream_t s) {
    if (!f || num_mis <= 0 || crop < 3 || crop > LFPM_MAX_CROP || max_neighbors < 1 ||
        max_neighbors > LFPM_MAX_NEIGHBORS || !(base > 0.0f))
        return set_error(LFPM_E_INVALID, "filter: bad parameters");
    if (f->max_triplet < 0 || f->max_triplet > 32)
        return set_error(LFPM_E_INVALID, "filter: max_triplet outside [0, 32]");
    if (!gray_dev || !centers_dev || !neighbor_ids_dev || !neighbor_counts_dev || !plane_in_dev ||
        !cost_in_dev || !disp_in_dev || !sel_in_dev || !plane_out_dev || !cost_out_dev ||
        !disp_out_dev || !sel_out_dev || !reason_dev)
        return set_error(LFPM_E_INVALID, "filter: null buffer");
    FilterArgs a;
    a.gray = gray_dev; a.centers = (const float2*)centers_dev; a.nbr = neighbor_ids_dev;
    a.cnt = neighbor_counts_dev; a.V = num_mis; a.W = crop; a.K = max_neighbors; a.base = base;
    a.f = *f;
    a.plane_in = (const float4*)plane_in_dev; a.cost_in = cost_in_dev;
    a.disp_in = (const float4*)disp_in_dev; a.sel_in = sel_in_dev;
    a.plane_out = (float4*)plane_out_dev; a.cost_out = cost_out_dev;
    a.disp_out = (float4*)disp_out_dev; a.sel_out = sel_out_dev; a.reason = reason_dev;
    a.mi_order = mi_order_dev; a.level_begin = 0;
    const int NT = 64 * ((crop * crop + 63) / 64);
    if (!ordered) {
        hipLaunchKernelGGL(k_filter<false>, dim3(8 * ((num_mis + 7) / 8)), dim3(NT), 0, s, a);
        return launch_check("filter");
    }
    // one launch per level, in order: the kernel boundary makes a level's outputs visible to
    // the next
    for (int l = 0; l < num_levels; ++l) {
        a.level_begin = level_start[l];
        const int n = level_start[l + 1] - level_start[l];
        hipLaunchKernelGGL(k_filter<true>, dim3(n), dim3(NT), 0, s, a);
    }
    return launch_check("filter_ordered");
}

int lfpm_filter_frame(const float* gray_dev, const float* centers_dev,
                      const int* neighbor_ids_dev, const int* neighbor_counts_dev, int num_mis,
                      int crop, int max_neighbors, float base, const lfpm_filter_params* f,
                      const float* plane_in_dev, const float* cost_in_dev,
                      const float* disp_in_dev, const uint32_t* sel_in_dev, float* plane_out_dev,
                      float* cost_out_dev, float* disp_out_dev, uint32_t* sel_out_dev,
                      uint8_t* reason_dev, void* stream) {
    return filter_impl(false, gray_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev,
                       num_mis, crop, max_neighbors, base, f, plane_in_dev, cost_in_dev,
                       disp_in_dev, sel_in_dev, plane_out_dev, cost_out_dev, disp_out_dev,
                       sel_out_dev, reason_dev, nullptr, nullptr, 0, (hipStream_t)stream);
}

int lfpm_filter_frame_ordered(const float* gray_dev, const float* centers_dev,
                              const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                              int num_mis, int crop, int max_neighbors, float base,
                              const lfpm_filter_params* f, const float* plane_in_dev,
                              const float* cost_in_dev, const float* disp_in_dev,
                              const uint32_t* sel_in_dev, float* plane_out_dev,
                              float* cost_out_dev, float* disp_out_dev, uint32_t* sel_out_dev,
                              uint8_t* reason_dev, const int* mi_order_dev,
                              const int* level_start, int num_levels, void* stream) {
    if (!mi_order_dev || !level_start || num_levels < 1 || num_levels > num_mis)
        return set_error(LFPM_E_INVALID, "filter_ordered: no level schedule");
    if (level_start[0] != 0 || level_start[num_levels] != num_mis)
        return set_error(LFPM_E_INVALID, "filter_ordered: the levels must cover MIs 0..%d",
                         num_mis - 1);
    for (int l = 0; l < num_levels; ++l)
        if (level_start[l + 1] <= level_start[l])
            return set_error(LFPM_E_INVALID, "filter_ordered: level %d is empty or reversed", l);
    if (plane_out_dev == plane_in_dev || cost_out_dev == cost_in_dev)
        return set_error(LFPM_E_INVALID, "filter_ordered: outputs must not alias the inputs");
    return filter_impl(true, gray_dev, centers_dev, neighbor_ids_dev, neighbor_counts_dev,
                       num_mis, crop, max_neighbors, base, f, plane_in_dev, cost_in_dev,
                       disp_in_dev, sel_in_dev, plane_out_dev, cost_out_dev, disp_out_dev,
                       sel_out_dev, reason_dev, mi_order_dev, level_start, num_levels,
                       (hipStream_t)stream);
}

int lfpm_refocus(const float* tiles_dev, const float* centers_dev, const int* lattice_to_mi_dev,
                 int lattice_cols, int lattice_rows, float pitch, int crop, int out_w, int out_h,
                 int num_slices, float vmin, float vmax, float* out_dev, void* stream) {
    if (!tiles_dev || !centers_dev || !lattice_to_mi_dev || !out_dev || lattice_cols <= 0 ||
        lattice_rows <= 0 || !(pitch > 0.0f) || crop < 2 || out_w <= 0 || out_h <= 0 ||
        num_slices <= 0 || !(vmin > 0.0f) || !(vmax >= vmin))
        return set_error(LFPM_E_INVALID, "refocus: bad arguments");
    // the tile kernel when every tile's candidate list fits: at most (rows spanned) x (columns
    // spanned) lattice cells for the largest reach
    const float vtop = std::max(vmin, vmax);
    const double reach = (double)vtop * (crop - 1) * 0.5 + 1.0;
    const double span = 16.0 + 2.0 * reach;
    const double nrows = std::ceil(span / (pitch * 0.8660254037844386)) + 2.0;
    const double ncand = nrows * (std::ceil(span / pitch) + 2.0);
    if (ncand <= kRefocusCands && nrows <= 32) {
        const int cap = (int)ncand;
        const size_t lds = (size_t)cap * (32 * sizeof(RefocusAxis) + sizeof(float2) + 8);
        const int ntiles = ((out_w + 15) / 16) * ((out_h + 15) / 16);
        dim3 grid(8 * ((ntiles + 7) / 8));
        hipLaunchKernelGGL(k_refocus_tile, grid, dim3(256), lds, (hipStream_t)stream, tiles_dev,
                           (const float2*)centers_dev, lattice_to_mi_dev, lattice_cols,
                           lattice_rows, pitch, crop, out_w, out_h, num_slices, vmin, vmax, cap,
                           out_dev);
        return launch_check("refocus");
    }
    dim3 grid((out_w + 15) / 16, (out_h + 15) / 16, num_slices);
    hipLaunchKernelGGL(k_refocus, grid, dim3(256), 0, (hipStream_t)stream, tiles_dev,
                       (const float2*)centers_dev, lattice_to_mi_dev, lattice_cols, lattice_rows,
                       pitch, crop, out_w, out_h, num_slices, vmin, vmax, out_dev);
    return launch_check("refocus");
}

int lfpm_fit_v2r(const float* a_l_dev, const float* v_dev, int n, double c_out[3], void* stream) {
    if (!a_l_dev || !v_dev || !c_out || n < 3)
        return set_error(LFPM_E_INVALID, "fit: need >= 3 samples");
    hipStream_t s = (hipStream_t)stream;
    const int blocks = std::min(1024, (n + 255) / 256);
    double* part = nullptr;
    HIP_TRY(hipMallocAsync((void**)&part, sizeof(double) * (blocks * 9 + 3), s));
    hipLaunchKernelGGL(k_fit_partial, dim3(blocks), dim3(256), 0, s, a_l_dev, v_dev, n, part);
    hipLaunchKernelGGL(k_fit_solve, dim3(1), dim3(64), 0, s, part, blocks, part + blocks * 9);
    int rc = launch_check("fit");
    if (rc == LFPM_OK) {
        HIP_TRY(hipMemcpyAsync(c_out, part + blocks * 9, sizeof(double) * 3,
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    HIP_TRY(hipFreeAsync(part, s));
    return rc;
}

int lfpm_set_work_counters(unsigned long long* counters_dev) {
    g_counters = counters_dev;
    return LFPM_OK;
}

int lfpm_v2r_convert(const float* v_dev, int n, const double c[3], double* z_dev, void* stream) {
    if (!v_dev || !z_dev || !c || n < 0) return set_error(LFPM_E_INVALID, "convert: bad args");
    if (n == 0) return LFPM_OK;
    const int blocks = std::min(65535, (n + 255) / 256);
    hipLaunchKernelGGL(k_v2r_convert, dim3(blocks), dim3(256), 0, (hipStream_t)stream, v_dev, n,
                       c[0], c[1], c[2], z_dev);
    return launch_check("v2r_convert");
}

} // extern "C"
