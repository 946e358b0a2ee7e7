// lfpm_products.hip — per-frame products of a solved frame, on the GPU:
//   * the raw-size disparity mosaic (every MI pixel whose cost passes the gate lands at its raw
//     position) and the frame's disparity range, and
//   * the virtual-depth (VD) map: every MI pixel splats f_v = mla_base / d at its projected raw
//     position; each raw pixel keeps the median of its samples, plus the "real colour" image,
//     gray or B, G, R (per channel: median colour, empty pixels filled by the directional
//     search, ten 5x5 median passes).
// Both are HBM-bound scatter / gather passes over V*crop^2 MI pixels and raw_w*raw_h raw
// pixels.  Every result is deterministic: scatters either resolve collisions by an order-free
// rule (highest MI index wins) or sort what they collected before reducing it.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cmath>

#include "../../include/lfpm.h"
#include "lfpm_internal.h"
#include "lfpm_median25.h"

namespace lfpm {
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanChunk = kScanThreads * kScanItems;

// Order-preserving float <-> int map, so float min/max can use integer atomics.
__device__ __forceinline__ int f2o(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float o2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

__global__ void k_fill_i32(int* __restrict__ p, long long n, int v) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// Range accumulators: kRangeSlots {ordered(min), ordered(max)} pairs, one per 128-byte line,
// so that the workgroups of a large grid spread their atomics over several addresses instead of
// serialising on one; k_range_out folds the slots.
constexpr int kRangeSlots = 64;
constexpr int kRangeStride = 32;  // ints: 128 bytes
constexpr size_t kRangeBytes = (size_t)kRangeSlots * kRangeStride * 4;

// slot (blockIdx.x % kRangeSlots) takes the min / max of the values a block saw; empty blocks
// add nothing.  k_range_init sets every slot to {ordered(+inf), ordered(-inf)}.
__device__ __forceinline__ void block_minmax(int* range, bool has, float val) {
    range += (blockIdx.x % kRangeSlots) * kRangeStride;
    __shared__ int smin, smax;
    if (threadIdx.x == 0) { smin = f2o(INFINITY); smax = f2o(-INFINITY); }
    __syncthreads();
    if (has) {
        atomicMin(&smin, f2o(val));
        atomicMax(&smax, f2o(val));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (smin != f2o(INFINITY)) atomicMin(&range[0], smin);
        if (smax != f2o(-INFINITY)) atomicMax(&range[1], smax);
    }
}

__global__ void k_range_init(int* range) {
    const int t = threadIdx.x;
    if (t < kRangeSlots) {
        range[t * kRangeStride] = f2o(INFINITY);
        range[t * kRangeStride + 1] = f2o(-INFINITY);
    }
}

__global__ void k_range_out(const int* __restrict__ range, float* __restrict__ out) {
    int lo = f2o(INFINITY), hi = f2o(-INFINITY);
    for (int t = 0; t < kRangeSlots; ++t) {
        lo = min(lo, range[t * kRangeStride]);
        hi = max(hi, range[t * kRangeStride + 1]);
    }
    out[0] = o2f(lo);
    out[1] = o2f(hi);
}

// Raw position of local pixel (x, y) of an MI centred at (cx, cy) in the mosaic: the crop's
// top-left corner c - (round(crop/2) - 1), truncated toward zero
// (DepthSolver::WriteDisMap_TileKey_new_AccuSequence, DepthSolver.cpp:6790-6812).
__device__ __forceinline__ int mosaic_origin(float c, int crop) {
#pragma clang fp contract(off)
    return (int)(c - (roundf((float)crop * 0.5f) - 1.0f));
}

// ------------------------------------------------------------------------------------------
// Disparity mosaic
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_mosaic_scatter(const float4* __restrict__ plane,
                                                        const float* __restrict__ cost,
                                                        const float2* __restrict__ centers,
                                                        int V, int crop, float max_cost,
                                                        int raw_w, int raw_h,
                                                        int* __restrict__ owner,
                                                        int* __restrict__ range) {
    const int WW = crop * crop;
    const long long M = (long long)V * WW;
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    float d = 0.0f;
    bool has = false;
    if (i < M) {
        const int v = (int)(i / WW), pix = (int)(i - (long long)v * WW);
        d = plane[i].w;
        has = d > 0.0f;
        if (has && cost[i] <= max_cost) {
            const float2 c = centers[v];
            const int X = mosaic_origin(c.x, crop) + pix % crop;
            const int Y = mosaic_origin(c.y, crop) + pix / crop;
            if (X >= 0 && X < raw_w && Y >= 0 && Y < raw_h)
                atomicMax(&owner[(long long)Y * raw_w + X], v);
        }
    }
    block_minmax(range, has, d);
}

__global__ void __launch_bounds__(256) k_mosaic_gather(const float4* __restrict__ plane,
                                                       const float2* __restrict__ centers,
                                                       const int* __restrict__ owner, int crop,
                                                       int raw_w, int raw_h,
                                                       float* __restrict__ out) {
    const long long N = (long long)raw_w * raw_h;
    const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (p >= N) return;
    const int o = owner[p];
    float d = 0.0f;
    if (o >= 0) {
        const int Y = (int)(p / raw_w), X = (int)(p - (long long)Y * raw_w);
        const float2 c = centers[o];
        const int x = X - mosaic_origin(c.x, crop), y = Y - mosaic_origin(c.y, crop);
        d = plane[(long long)o * crop * crop + y * crop + x].w;
    }
    out[p] = d;
}

// ------------------------------------------------------------------------------------------
// Virtual-depth splat (DepthSolver::Virtual_depth_map_TileKeySequence, DepthSolver.cpp:7832-7905)
// ------------------------------------------------------------------------------------------
// Raw target of MI pixel i (or -1) and its virtual depth f_v.
__device__ __forceinline__ long long vd_target(const float4* __restrict__ plane,
                                               const float2* __restrict__ centers, long long i,
                                               int crop, float mla_base, int raw_w, int raw_h,
                                               float& fv) {
#pragma clang fp contract(off)
    const int WW = crop * crop;
    const int v = (int)(i / WW), pix = (int)(i - (long long)v * WW);
    const int x = pix % crop, y = pix / crop;
    float d = plane[i].w;
    if (d <= 0.0f) d = 1.0f;
    fv = mla_base / d;
    const float half1 = (float)(crop - 1) * 0.5f + 1.0f;
    const int xm = (int)((float)x - half1), ym = (int)((float)y - half1);
    const float2 c = centers[v];
    const float Xv = (float)xm * fv + c.x, Yv = (float)ym * fv + c.y;
    const int X = (int)roundf(Xv), Y = (int)roundf(Yv);
    if (X > 0 && X < raw_w && Y > 0 && Y < raw_h) return (long long)Y * raw_w + X;
    return -1;
}

__global__ void __launch_bounds__(256) k_vd_count(const float4* __restrict__ plane,
                                                  const float2* __restrict__ centers, int V,
                                                  int crop, float mla_base, int raw_w, int raw_h,
                                                  int* __restrict__ count) {
    const long long M = (long long)V * crop * crop;
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= M) return;
    float fv;
    const long long t = vd_target(plane, centers, i, crop, mla_base, raw_w, raw_h, fv);
    if (t >= 0) atomicAdd(&count[t], 1);
}

// Exclusive scan of n ints (three passes: chunk sums, scan of sums, chunk scans).
__global__ void __launch_bounds__(kScanThreads) k_scan_sums(const int* __restrict__ in, long long n,
                                                            int* __restrict__ sums) {
    __shared__ int s[kScanThreads];
    const long long base = (long long)blockIdx.x * kScanChunk;
    int acc = 0;
    for (int k = 0; k < kScanItems; ++k) {
        const long long i = base + (long long)k * kScanThreads + threadIdx.x;
        if (i < n) acc += in[i];
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kScanThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = s[0];
}

// Inclusive Hillis-Steele scan of one value per thread over the block; returns the exclusive
// prefix of this thread and the block total.
__device__ __forceinline__ int block_exclusive(int v, int& total) {
    __shared__ int s[kScanThreads];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int add = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    total = s[kScanThreads - 1];
    const int incl = s[threadIdx.x];
    __syncthreads();
    return incl - v;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_sums_scan(int* __restrict__ sums, int nb) {
    // one block: each thread owns a contiguous run of ceil(nb / threads) sums
    const int per = (nb + kScanThreads - 1) / kScanThreads;
    const int b0 = threadIdx.x * per;
    int acc = 0;
    for (int k = 0; k < per; ++k)
        if (b0 + k < nb) acc += sums[b0 + k];
    int total;
    int run = block_exclusive(acc, total);
    for (int k = 0; k < per; ++k)
        if (b0 + k < nb) {
            const int v = sums[b0 + k];
            sums[b0 + k] = run;
            run += v;
        }
}

__global__ void __launch_bounds__(kScanThreads) k_scan_chunks(const int* __restrict__ in,
                                                              long long n,
                                                              const int* __restrict__ sums,
                                                              int* __restrict__ out) {
    // thread t owns items [base + t*kScanItems, +kScanItems) (contiguous per thread)
    const long long base = (long long)blockIdx.x * kScanChunk + (long long)threadIdx.x * kScanItems;
    int vals[kScanItems];
    int acc = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const long long i = base + k;
        vals[k] = i < n ? in[i] : 0;
        acc += vals[k];
    }
    int total;
    int run = block_exclusive(acc, total) + sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const long long i = base + k;
        if (i < n) out[i] = run;
        run += vals[k];
    }
}

// Samples: sv = f_v, sc = colour; with colour crops (bgr: V*crop^2 pixels B, G, R) sc holds
// three planes of M bytes (channel c of sample slot s at sc[c*M + s]).
__global__ void __launch_bounds__(256) k_vd_scatter(const float4* __restrict__ plane,
                                                    const float2* __restrict__ centers,
                                                    const float* __restrict__ gray,
                                                    const uint8_t* __restrict__ bgr, int V,
                                                    int crop, float mla_base, int raw_w,
                                                    int raw_h, int* __restrict__ count,
                                                    const int* __restrict__ offs,
                                                    float* __restrict__ sv,
                                                    uint8_t* __restrict__ sc) {
    const long long M = (long long)V * crop * crop;
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= M) return;
    float fv;
    const long long t = vd_target(plane, centers, i, crop, mla_base, raw_w, raw_h, fv);
    if (t < 0) return;
    const int slot = offs[t] + atomicSub(&count[t], 1) - 1;
    sv[slot] = fv;
    if (bgr) {
        sc[slot] = bgr[3 * i];
        sc[M + slot] = bgr[3 * i + 1];
        sc[2 * M + slot] = bgr[3 * i + 2];
    } else {
        sc[slot] = gray ? (uint8_t)gray[i] : (uint8_t)0;
    }
}

// Median of each raw pixel's samples (sorted in place).  VD value: the reference's index rule
// (s[n/2+1] + s[(n+1)/2]) / 2, with indices clamped to n-1 (the rule reads past the end for
// n <= 2); colour, each of the CH channels sorted on its own (color0/1/2 of
// DepthSolver.cpp:7919-7927): (c[n/2] + c[(n+1)/2-1]) / 2 in integers.  Colour planes: sample
// planes sc + c*M, median planes col + c*N.
template <int CH>
__global__ void __launch_bounds__(256) k_vd_median(const int* __restrict__ offs, long long N,
                                                   long long M, float* __restrict__ sv,
                                                   uint8_t* __restrict__ sc,
                                                   float* __restrict__ vd,
                                                   uint8_t* __restrict__ col,
                                                   int* __restrict__ range) {
    const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    bool has = false;
    float m = 0.0f;
    if (p < N) {
        const int b = offs[p], n = offs[p + 1] - b;
        uint8_t cm[CH];
#pragma unroll
        for (int ch = 0; ch < CH; ++ch) cm[ch] = 0;
        if (n == 1 || n == 2) {
            // the index rule reads s[n-1] twice for n <= 2 (the largest sample; halving the
            // doubled value is exact), the colour rule averages the two samples
            const float s0 = sv[b], s1 = n == 2 ? sv[b + 1] : s0;
            m = ((s0 > s1 ? s0 : s1) + (s0 > s1 ? s0 : s1)) / 2.0f;
#pragma unroll
            for (int ch = 0; ch < CH; ++ch) {
                const uint8_t* c = sc + ch * M;
                const int c0 = c[b], c1 = n == 2 ? c[b + 1] : c0;
                cm[ch] = (uint8_t)((c0 + c1) / 2);
            }
            has = true;
        } else if (n <= 8 && n > 0) {
            // 8-wire sorting network in registers, padding above every sample (+inf, 256)
            float f[8];
            int q[CH][8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                f[k] = k < n ? sv[b + k] : INFINITY;
#pragma unroll
                for (int ch = 0; ch < CH; ++ch) q[ch][k] = k < n ? (int)sc[ch * M + b + k] : 256;
            }
            auto cx = [&](int i, int j) {
                const float lo = fminf(f[i], f[j]), hi = fmaxf(f[i], f[j]);
                f[i] = lo;
                f[j] = hi;
#pragma unroll
                for (int ch = 0; ch < CH; ++ch) {
                    const int a = min(q[ch][i], q[ch][j]), c = max(q[ch][i], q[ch][j]);
                    q[ch][i] = a;
                    q[ch][j] = c;
                }
            };
            // Batcher's odd-even merge sort, 8 inputs (19 comparators)
            cx(0, 1); cx(2, 3); cx(4, 5); cx(6, 7);
            cx(0, 2); cx(1, 3); cx(4, 6); cx(5, 7);
            cx(1, 2); cx(5, 6);
            cx(0, 4); cx(1, 5); cx(2, 6); cx(3, 7);
            cx(2, 4); cx(3, 5);
            cx(1, 2); cx(3, 4); cx(5, 6);
            const int i1 = min(n / 2 + 1, n - 1), i2 = min((n + 1) / 2, n - 1);
            const int j1 = n / 2, j2 = (n + 1) / 2 - 1;
            float a1 = f[0], a2 = f[0];
            int d1[CH], d2[CH];
#pragma unroll
            for (int ch = 0; ch < CH; ++ch) d1[ch] = d2[ch] = q[ch][0];
#pragma unroll
            for (int k = 1; k < 8; ++k) {
                a1 = k == i1 ? f[k] : a1;
                a2 = k == i2 ? f[k] : a2;
#pragma unroll
                for (int ch = 0; ch < CH; ++ch) {
                    d1[ch] = k == j1 ? q[ch][k] : d1[ch];
                    d2[ch] = k == j2 ? q[ch][k] : d2[ch];
                }
            }
            m = (a1 + a2) / 2.0f;
#pragma unroll
            for (int ch = 0; ch < CH; ++ch) cm[ch] = (uint8_t)((d1[ch] + d2[ch]) / 2);
            has = true;
        } else if (n > 0) {
            float* s = sv + b;
            for (int a = 1; a < n; ++a) {
                const float x = s[a];
                int j = a - 1;
                while (j >= 0 && s[j] > x) { s[j + 1] = s[j]; --j; }
                s[j + 1] = x;
            }
            const int i1 = min(n / 2 + 1, n - 1), i2 = min((n + 1) / 2, n - 1);
            m = (s[i1] + s[i2]) / 2.0f;
            for (int ch = 0; ch < CH; ++ch) {
                uint8_t* c = sc + ch * M + b;
                for (int a = 1; a < n; ++a) {
                    const uint8_t y = c[a];
                    int j = a - 1;
                    while (j >= 0 && c[j] > y) { c[j + 1] = c[j]; --j; }
                    c[j + 1] = y;
                }
                cm[ch] = (uint8_t)(((int)c[n / 2] + (int)c[(n + 1) / 2 - 1]) / 2);
            }
            has = true;
        }
        vd[p] = m;
#pragma unroll
        for (int ch = 0; ch < CH; ++ch) col[ch * N + p] = cm[ch];
    }
    block_minmax(range, has, m);
}

// Search tables of the colour fill (thread per raw row): in-row next / previous pixel with
// samples, and each row's first / last such pixel.
// Row links: nr[p] / pr[p] = the next / previous raw pixel with samples in p's row (-1 if
// none), rfirst / rlast = the row's first / last such pixel.  One workgroup per row: each lane
// scans a contiguous segment; the carries across segments come from a workgroup min / max scan.
constexpr int kFillThreads = 256;
constexpr int kFillSeg = 64;  // rows (columns) or steps (diagonals) per segment

__global__ void __launch_bounds__(kFillThreads) k_fill_rows(const int* __restrict__ offs,
                                                            int raw_w, int* __restrict__ nr,
                                                            int* __restrict__ pr,
                                                            int* __restrict__ rfirst,
                                                            int* __restrict__ rlast) {
    __shared__ int sf[kFillThreads], sl[kFillThreads];
    const int y = blockIdx.x, t = threadIdx.x;
    const long long row = (long long)y * raw_w;
    const int per = (raw_w + kFillThreads - 1) / kFillThreads;
    const int x0 = min(t * per, raw_w), x1 = min(x0 + per, raw_w);
    int lf = INT_MAX, ll = -1;
    for (int x = x0; x < x1; ++x)
        if (offs[row + x + 1] > offs[row + x]) {
            if (lf == INT_MAX) lf = (int)(row + x);
            ll = (int)(row + x);
        }
    sf[t] = lf;
    sl[t] = ll;
    __syncthreads();
    // inclusive scans: sl = max over lanes <= t, sf = min over lanes >= t
    for (int o = 1; o < kFillThreads; o <<= 1) {
        const int a = t >= o ? sl[t - o] : -1;
        const int b = t + o < kFillThreads ? sf[t + o] : INT_MAX;
        __syncthreads();
        sl[t] = max(sl[t], a);
        sf[t] = min(sf[t], b);
        __syncthreads();
    }
    int prv = t > 0 ? sl[t - 1] : -1;
    for (int x = x0; x < x1; ++x) {
        pr[row + x] = prv;
        if (offs[row + x + 1] > offs[row + x]) prv = (int)(row + x);
    }
    int nxt = t + 1 < kFillThreads ? sf[t + 1] : INT_MAX;
    for (int x = x1 - 1; x >= x0; --x) {
        nr[row + x] = nxt == INT_MAX ? -1 : nxt;
        if (offs[row + x + 1] > offs[row + x]) nxt = (int)(row + x);
    }
    if (t == 0) {
        rfirst[y] = sf[0] == INT_MAX ? -1 : sf[0];
        rlast[y] = sl[kFillThreads - 1];
    }
}


// nextRowFirst[y] = first sampled pixel in a row below y (N if none); prevRowLast[y] = last
// sampled pixel in a row above y (-1 if none).  One thread: raw_h steps.
__global__ void __launch_bounds__(kFillThreads) k_fill_rowlinks(const int* __restrict__ rfirst,
                                                                const int* __restrict__ rlast,
                                                                int raw_h, long long N,
                                                                int* __restrict__ nrf,
                                                                int* __restrict__ prl) {
    // nrf[y] = the first pixel with samples in the rows after y (N if none), prl[y] = the last
    // one in the rows before y (-1 if none): exclusive suffix-min / prefix-max over rows, one
    // workgroup, each lane a contiguous run of rows
    __shared__ int sf[kFillThreads], sl[kFillThreads];
    const int t = threadIdx.x;
    const int per = (raw_h + kFillThreads - 1) / kFillThreads;
    const int y0 = min(t * per, raw_h), y1 = min(y0 + per, raw_h);
    int lf = INT_MAX, ll = -1;
    for (int y = y0; y < y1; ++y) {
        if (rfirst[y] >= 0 && lf == INT_MAX) lf = rfirst[y];
        if (rlast[y] >= 0) ll = rlast[y];
    }
    sf[t] = lf;
    sl[t] = ll;
    __syncthreads();
    for (int o = 1; o < kFillThreads; o <<= 1) {
        const int a = t >= o ? sl[t - o] : -1;
        const int b = t + o < kFillThreads ? sf[t + o] : INT_MAX;
        __syncthreads();
        sl[t] = max(sl[t], a);
        sf[t] = min(sf[t], b);
        __syncthreads();
    }
    int prv = t > 0 ? sl[t - 1] : -1;
    for (int y = y0; y < y1; ++y) {
        prl[y] = prv;
        if (rlast[y] >= 0) prv = rlast[y];
    }
    int nxt = t + 1 < kFillThreads ? sf[t + 1] : INT_MAX;
    for (int y = y1 - 1; y >= y0; --y) {
        nrf[y] = nxt == INT_MAX ? (int)N : nxt;
        if (rfirst[y] >= 0) nxt = rfirst[y];
    }
}

// Chains of the down-left step (linear stride raw_w - 1; thread per chain): next sampled index
// strictly after p along the chain, N if none.
// Anti-diagonal links: nd[p] = the next pixel with samples on p's chain p, p + (raw_w - 1), ...
// (N if none).  Chains are cut into kFillSeg-step segments (lane = chain, so a step reads
// consecutive pixels across lanes): the first pass records each segment's first pixel with
// samples, the second turns those into carries from the right, the third writes the links.
__device__ __forceinline__ long long diag_len(long long c, long long step, long long N) {
    return (N - 1 - c) / step + 1;
}

__global__ void k_diag_seg(const int* __restrict__ offs, long long step, long long N,
                           int nseg, int* __restrict__ segv) {
    const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const int sg = blockIdx.y;
    if (c >= step) return;
    const long long len = diag_len(c, step, N);
    const long long k0 = (long long)sg * kFillSeg, k1 = min(k0 + kFillSeg, len);
    int first = (int)N;
    for (long long k = k1 - 1; k >= k0; --k) {
        const long long p = c + k * step;
        if (offs[p + 1] > offs[p]) first = (int)p;
    }
    segv[(long long)sg * step + c] = first;
}

__global__ void k_diag_carry(long long step, int nseg, int N, int* __restrict__ segv) {
    const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (c >= step) return;
    int carry = N;  // segment sg gets the first pixel with samples in the segments after it
    for (int sg = nseg - 1; sg >= 0; --sg) {
        int* q = segv + (long long)sg * step + c;
        const int v = *q;
        *q = carry;
        if (v < N) carry = v;
    }
}

__global__ void k_diag_links(const int* __restrict__ offs, long long step, long long N,
                             const int* __restrict__ segv, int* __restrict__ nd) {
    const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const int sg = blockIdx.y;
    if (c >= step) return;
    const long long len = diag_len(c, step, N);
    const long long k0 = (long long)sg * kFillSeg, k1 = min(k0 + kFillSeg, len);
    int nxt = segv[(long long)sg * step + c];
    for (long long k = k1 - 1; k >= k0; --k) {
        const long long p = c + k * step;
        nd[p] = nxt;
        if (offs[p + 1] > offs[p]) nxt = (int)p;
    }
}


// Columns (thread per column): nearest sampled pixel above p in its column, -1 if none.
// Column links: pu[p] = the previous pixel with samples in p's column (-1 if none); the same
// three passes over kFillSeg-row segments (lane = column), carries from the top.
__global__ void k_cols_seg(const int* __restrict__ offs, int raw_w, int raw_h,
                           int* __restrict__ segv) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, sg = blockIdx.y;
    if (x >= raw_w) return;
    const int y0 = sg * kFillSeg, y1 = min(y0 + kFillSeg, raw_h);
    int last = -1;
    for (int y = y0; y < y1; ++y) {
        const long long p = (long long)y * raw_w + x;
        if (offs[p + 1] > offs[p]) last = (int)p;
    }
    segv[(long long)sg * raw_w + x] = last;
}

__global__ void k_cols_carry(int raw_w, int nseg, int* __restrict__ segv) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= raw_w) return;
    int carry = -1;  // segment sg gets the last pixel with samples in the segments above it
    for (int sg = 0; sg < nseg; ++sg) {
        int* q = segv + (long long)sg * raw_w + x;
        const int v = *q;
        *q = carry;
        if (v >= 0) carry = v;
    }
}

__global__ void k_cols_links(const int* __restrict__ offs, int raw_w, int raw_h,
                             const int* __restrict__ segv, int* __restrict__ pu) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, sg = blockIdx.y;
    if (x >= raw_w) return;
    const int y0 = sg * kFillSeg, y1 = min(y0 + kFillSeg, raw_h);
    int prv = segv[(long long)sg * raw_w + x];
    for (int y = y0; y < y1; ++y) {
        const long long p = (long long)y * raw_w + x;
        pu[p] = prv;
        if (offs[p + 1] > offs[p]) prv = (int)p;
    }
}


// Colour of pixels without samples: the reference probes, for s = 1, 2, ..., the linear indices
// p + s, p - s, p + s*(raw_w-1), p - s*raw_w in that order, stops at the first probe that leaves
// [0, N) (colour stays 0) or hits a pixel with samples (takes its median colour)
// (DepthSolver.cpp:7955-7993).  Here each probe sequence's first hit and first exit are read
// from the tables, and the earliest event in probe order decides.
__global__ void __launch_bounds__(256) k_fill_resolve(const int* __restrict__ offs, int raw_w,
                                                      int raw_h, const int* __restrict__ nr,
                                                      const int* __restrict__ pr,
                                                      const int* __restrict__ nrf,
                                                      const int* __restrict__ prl,
                                                      const int* __restrict__ nd,
                                                      const int* __restrict__ pu,
                                                      const uint8_t* __restrict__ colm,
                                                      int nch, uint8_t* __restrict__ out) {
    // colm and out: nch planes of N bytes; every channel takes the colour of the same pixel
    const long long N = (long long)raw_w * raw_h;
    const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (p >= N) return;
    if (offs[p + 1] > offs[p]) {
        for (int ch = 0; ch < nch; ++ch) out[ch * N + p] = colm[ch * N + p];
        return;
    }
    const int y = (int)(p / raw_w);
    const long long INF = 0x7fffffffffffffffLL;
    long long best_t = INF;
    long long best_k = -1;
    auto consider = [&](int c, long long s_hit, long long k_hit, long long s_exit) {
        const bool hit = s_hit < s_exit;
        const long long s = hit ? s_hit : s_exit;
        if (s == INF) return;
        const long long t = 4 * s + c;
        if (t < best_t) { best_t = t; best_k = hit ? k_hit : -1; }
    };
    {   // p + s
        const long long k = nr[p] >= 0 ? nr[p] : nrf[y];
        consider(0, k < N ? k - p : INF, k, N - p);
    }
    {   // p - s
        const long long k = pr[p] >= 0 ? pr[p] : prl[y];
        consider(1, k >= 0 ? p - k : INF, k, p + 1);
    }
    if (raw_w >= 2) {   // p + s*(raw_w - 1)
        const long long st = raw_w - 1;
        const long long k = nd[p];
        consider(2, k < N ? (k - p) / st : INF, k, (N - p + st - 1) / st);
    }
    {   // p - s*raw_w
        const long long k = pu[p];
        consider(3, k >= 0 ? (p - k) / raw_w : INF, k, (long long)y + 1);
    }
    for (int ch = 0; ch < nch; ++ch) out[ch * N + p] = best_k >= 0 ? colm[ch * N + best_k] : (uint8_t)0;
}

// 5x5 median with replicated borders (cv::medianBlur's documented border rule).  One lane per
// two horizontally adjacent pixels: their 25 window values are packed into the 16-bit halves of
// one word, and the median-of-25 selection network runs on packed 16-bit min / max.
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// out: pixel (x, y) at out[(y * raw_w + x) * ostride] (ostride 3: one channel of an
// interleaved B, G, R image).
__global__ void __launch_bounds__(256) k_median5(const uint8_t* __restrict__ in, int raw_w,
                                                 int raw_h, uint8_t* __restrict__ out,
                                                 int ostride) {
    const int pairs = (raw_w + 1) >> 1;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)pairs * raw_h) return;
    const int y = (int)(t / pairs);
    const int x0 = (int)(t - (long long)y * pairs) * 2;
    int cx[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) cx[i] = min(max(x0 + i - 2, 0), raw_w - 1);
    u16x2 v[25];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        const uint8_t* row = in + (long long)min(max(y + j - 2, 0), raw_h - 1) * raw_w;
        unsigned short c[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) c[i] = row[cx[i]];
#pragma unroll
        for (int i = 0; i < 5; ++i) v[j * 5 + i] = u16x2{c[i], c[i + 1]};
    }
    const u16x2 m = median25(
        v, [](u16x2 a, u16x2 b) { return __builtin_elementwise_min(a, b); },
        [](u16x2 a, u16x2 b) { return __builtin_elementwise_max(a, b); });
    uint8_t* o = out + ((long long)y * raw_w + x0) * ostride;
    o[0] = (uint8_t)m.x;
    if (x0 + 1 < raw_w) o[ostride] = (uint8_t)m.y;
}

int grid_for(long long n, int threads) {
    long long g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > 0x7fffffffLL) g = 0x7fffffffLL;
    return (int)g;
}

size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

// VD point cloud (DepthSolver::Virtual_depth_map_TileKeySequence, DepthSolver.cpp:8019-8053) as
// the binary payload of StoreColorPlyFileBinaryPointCloud (MVStereo/ACMP.cpp:451-500): one
// 27-byte record {x, y, z, nx, ny, nz (float32), red, green, blue (uint8)} per raw pixel with a
// VD sample, in row-major pixel order.  Flag pass, exclusive scan, emit pass.
__global__ void __launch_bounds__(256) k_pt_flag(const float* __restrict__ vd, long long N,
                                                 int* __restrict__ flag) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i > N) return;
    flag[i] = (i < N && vd[i] > 0.0f) ? 1 : 0;
}

__device__ __forceinline__ bool ply_open(float v) { return v < FLT_MAX && v > -FLT_MAX; }

__global__ void __launch_bounds__(256) k_pt_emit(const float* __restrict__ vd,
                                                 const uint8_t* __restrict__ color, int nch,
                                                 const int* __restrict__ offs, int raw_w,
                                                 long long N, float ox, float oy, float pix,
                                                 float f, float B, float bl0,
                                                 uint8_t* __restrict__ rec,
                                                 unsigned* __restrict__ count) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i == 0) *count = (unsigned)offs[N];
    if (i >= N) return;
    const float vdm = vd[i];
    if (!(vdm > 0.0f)) return;
    const int y = (int)(i / raw_w), x = (int)(i - (long long)y * raw_w);
    // the reference's expression order: image-plane point in FP64, the virtual depth mapped by
    // the FP32 line vdm * B + bl0 (two roundings, no contraction), then the thin-lens map
    const double tx = ((double)x - (double)ox) * (double)pix;
    const double ty = ((double)y - (double)oy) * (double)pix;
    float tzf;
    {
#pragma clang fp contract(off)
        tzf = vdm * B + bl0;
    }
    const double tz = (double)tzf;
    const double fd = (double)f;
    float X = (float)(fd / (tz - fd) * tz);
    float Xx = (float)(-fd / (tz - fd) * tx);
    float Xy = (float)(-fd / (tz - fd) * ty);
    // the writer's finiteness test (z admits -FLT_MAX, x and y do not)
    if (!ply_open(Xx) || !ply_open(Xy) || !(X < FLT_MAX && X >= -FLT_MAX)) Xx = Xy = X = 0.0f;
    uint8_t r[27];
    const float v[6] = {Xx, Xy, X, 0.0f, 0.0f, 0.0f};
    __builtin_memcpy(r, v, 24);
    if (nch == 3) {
        // red, green, blue of a B, G, R image: the writer emits color.z, .y, .x
        r[24] = color[3 * i + 2];
        r[25] = color[3 * i + 1];
        r[26] = color[3 * i];
    } else {
        const uint8_t g = color ? color[i] : (uint8_t)0;  // the gray colour image
        r[24] = g;
        r[25] = g;
        r[26] = g;
    }
    uint8_t* dst = rec + (size_t)offs[i] * 27;
#pragma unroll
    for (int k = 0; k < 27; ++k) dst[k] = r[k];
}

struct VdLayout {
    size_t count, offs, sums, sv, sc, nr, pr, nd, pu, rfirst, rlast, nrf, prl, col, range, segc,
        segd, total;
};

VdLayout vd_layout(int V, int crop, int raw_w, int raw_h) {
    const size_t N = (size_t)raw_w * raw_h, M = (size_t)V * crop * crop;
    const size_t nb = (N + 1 + kScanChunk - 1) / kScanChunk;
    VdLayout L;
    size_t o = 0;
    L.count = o; o += al((N + 1) * 4);
    L.offs = o; o += al((N + 1) * 4);
    L.sums = o; o += al(nb * 4);
    L.sv = o; o += al(M * 4);
    L.sc = o; o += al(3 * M);
    L.nr = o; o += al(N * 4);
    L.pr = o; o += al(N * 4);
    L.nd = o; o += al(N * 4);
    L.pu = o; o += al(N * 4);
    L.rfirst = o; o += al((size_t)raw_h * 4);
    L.rlast = o; o += al((size_t)raw_h * 4);
    L.nrf = o; o += al((size_t)raw_h * 4);
    L.prl = o; o += al((size_t)raw_h * 4);
    L.col = o; o += al(3 * N);
    L.range = o; o += al(kRangeBytes);
    // column segments (raw_w per segment) and diagonal segments (raw_w - 1 chains per segment)
    const size_t ncs = ((size_t)raw_h + kFillSeg - 1) / kFillSeg;
    const size_t step = raw_w > 1 ? (size_t)raw_w - 1 : 1;
    const size_t nds = ((N - 1) / step + 1 + kFillSeg - 1) / kFillSeg;
    L.segc = o; o += al(ncs * (size_t)raw_w * 4);
    L.segd = o; o += al(nds * step * 4);
    L.total = o;
    return L;
}

size_t mosaic_bytes(int raw_w, int raw_h) { return al((size_t)raw_w * raw_h * 4) + al(kRangeBytes); }

int check_launch(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(LFPM_E_HIP, "%s: %s", what, hipGetErrorString(e));
    return LFPM_OK;
}

int check_frame(int V, int crop, int raw_w, int raw_h) {
    if (V <= 0 || crop <= 0 || crop > LFPM_MAX_CROP || raw_w <= 0 || raw_h <= 0)
        return set_error(LFPM_E_INVALID, "products: bad sizes (V=%d crop=%d raw %dx%d)", V, crop,
                         raw_w, raw_h);
    if ((long long)raw_w * raw_h + 1 >= 0x7fffffffLL || (long long)V * crop * crop >= 0x7fffffffLL)
        return set_error(LFPM_E_INVALID, "products: frame too large for 32-bit indices");
    return LFPM_OK;
}

}  // namespace
}  // namespace lfpm

using namespace lfpm;

extern "C" {

size_t lfpm_products_workspace_size(int num_mis, int crop, int raw_w, int raw_h) {
    if (num_mis <= 0 || crop <= 0 || raw_w <= 0 || raw_h <= 0) return 0;
    const size_t a = mosaic_bytes(raw_w, raw_h), b = vd_layout(num_mis, crop, raw_w, raw_h).total;
    return a > b ? a : b;
}

int lfpm_disparity_mosaic(const float* plane_dev, const float* cost_dev, const float* centers_dev,
                          int num_mis, int crop, float max_cost, int raw_w, int raw_h,
                          float* mosaic_dev, float* range_dev, void* workspace_dev,
                          size_t workspace_bytes, void* stream) {
    if (int e = check_frame(num_mis, crop, raw_w, raw_h)) return e;
    if (!plane_dev || !cost_dev || !centers_dev || !mosaic_dev || !workspace_dev)
        return set_error(LFPM_E_INVALID, "lfpm_disparity_mosaic: null pointer");
    if (workspace_bytes < mosaic_bytes(raw_w, raw_h))
        return set_error(LFPM_E_INVALID, "lfpm_disparity_mosaic: workspace %zu < %zu bytes",
                         workspace_bytes, mosaic_bytes(raw_w, raw_h));
    hipStream_t s = (hipStream_t)stream;
    const long long N = (long long)raw_w * raw_h, M = (long long)num_mis * crop * crop;
    unsigned char* ws = (unsigned char*)workspace_dev;
    int* owner = (int*)ws;
    int* range = (int*)(ws + al((size_t)N * 4));
    k_fill_i32<<<grid_for(N, 256 * 8), 256, 0, s>>>(owner, N, -1);
    k_range_init<<<1, kRangeSlots, 0, s>>>(range);
    k_mosaic_scatter<<<grid_for(M, 256), 256, 0, s>>>(
        (const float4*)plane_dev, cost_dev, (const float2*)centers_dev, num_mis, crop, max_cost,
        raw_w, raw_h, owner, range);
    k_mosaic_gather<<<grid_for(N, 256), 256, 0, s>>>((const float4*)plane_dev,
                                                      (const float2*)centers_dev, owner, crop,
                                                      raw_w, raw_h, mosaic_dev);
    if (range_dev) k_range_out<<<1, 1, 0, s>>>(range, range_dev);
    return check_launch("lfpm_disparity_mosaic");
}

}  // extern "C"

namespace lfpm {
namespace {
// The VD map with its colour image from the gray crops (gray_dev, one channel) or the colour
// crops (bgr_dev, B, G, R: color_dev is then raw_h x raw_w x 3, interleaved).
int vd_map_impl(const float* plane_dev, const float* centers_dev, const float* gray_dev,
                const uint8_t* bgr_dev, int num_mis, int crop, float mla_base, int raw_w,
                int raw_h, float* vd_dev, uint8_t* color_dev, float* range_dev,
                void* workspace_dev, size_t workspace_bytes, void* stream) {
    if (int e = check_frame(num_mis, crop, raw_w, raw_h)) return e;
    if (!plane_dev || !centers_dev || !vd_dev || !workspace_dev)
        return set_error(LFPM_E_INVALID, "lfpm_virtual_depth_map: null pointer");
    if (!(mla_base > 0.0f))
        return set_error(LFPM_E_INVALID, "lfpm_virtual_depth_map: mla_base must be > 0");
    const VdLayout L = vd_layout(num_mis, crop, raw_w, raw_h);
    if (workspace_bytes < L.total)
        return set_error(LFPM_E_INVALID, "lfpm_virtual_depth_map: workspace %zu < %zu bytes",
                         workspace_bytes, L.total);
    hipStream_t s = (hipStream_t)stream;
    const long long N = (long long)raw_w * raw_h, M = (long long)num_mis * crop * crop;
    unsigned char* ws = (unsigned char*)workspace_dev;
    int* count = (int*)(ws + L.count);
    int* offs = (int*)(ws + L.offs);
    int* sums = (int*)(ws + L.sums);
    float* sv = (float*)(ws + L.sv);
    uint8_t* sc = ws + L.sc;
    int* range = (int*)(ws + L.range);
    const float4* plane = (const float4*)plane_dev;
    const float2* centers = (const float2*)centers_dev;
    const int nb = (int)((N + 1 + kScanChunk - 1) / kScanChunk);

    k_fill_i32<<<grid_for(N + 1, 256 * 8), 256, 0, s>>>(count, N + 1, 0);
    k_range_init<<<1, kRangeSlots, 0, s>>>(range);
    k_vd_count<<<grid_for(M, 256), 256, 0, s>>>(plane, centers, num_mis, crop, mla_base, raw_w,
                                                 raw_h, count);
    k_scan_sums<<<nb, kScanThreads, 0, s>>>(count, N + 1, sums);
    k_scan_sums_scan<<<1, kScanThreads, 0, s>>>(sums, nb);
    k_scan_chunks<<<nb, kScanThreads, 0, s>>>(count, N + 1, sums, offs);
    k_vd_scatter<<<grid_for(M, 256), 256, 0, s>>>(plane, centers, gray_dev, bgr_dev, num_mis,
                                                   crop, mla_base, raw_w, raw_h, count, offs, sv,
                                                   sc);
    uint8_t* colm = ws + L.col;
    const int nch = bgr_dev ? 3 : 1;
    if (bgr_dev)
        k_vd_median<3><<<grid_for(N, 256), 256, 0, s>>>(offs, N, M, sv, sc, vd_dev, colm, range);
    else
        k_vd_median<1><<<grid_for(N, 256), 256, 0, s>>>(offs, N, M, sv, sc, vd_dev, colm, range);
    if (range_dev) k_range_out<<<1, 1, 0, s>>>(range, range_dev);
    if (color_dev) {
        int* nr = (int*)(ws + L.nr);
        int* pr = (int*)(ws + L.pr);
        int* nd = (int*)(ws + L.nd);
        int* pu = (int*)(ws + L.pu);
        int* rfirst = (int*)(ws + L.rfirst);
        int* rlast = (int*)(ws + L.rlast);
        int* nrf = (int*)(ws + L.nrf);
        int* prl = (int*)(ws + L.prl);
        k_fill_rows<<<raw_h, kFillThreads, 0, s>>>(offs, raw_w, nr, pr, rfirst, rlast);
        k_fill_rowlinks<<<1, kFillThreads, 0, s>>>(rfirst, rlast, raw_h, N, nrf, prl);
        if (raw_w >= 2) {
            const long long step = raw_w - 1;
            const int nds = (int)(((N - 1) / step + 1 + kFillSeg - 1) / kFillSeg);
            int* segd = (int*)(ws + L.segd);
            const dim3 g((unsigned)grid_for(step, 256), (unsigned)nds);
            k_diag_seg<<<g, 256, 0, s>>>(offs, step, N, nds, segd);
            k_diag_carry<<<grid_for(step, 256), 256, 0, s>>>(step, nds, (int)N, segd);
            k_diag_links<<<g, 256, 0, s>>>(offs, step, N, segd, nd);
        }
        {
            const int ncs = (raw_h + kFillSeg - 1) / kFillSeg;
            int* segc = (int*)(ws + L.segc);
            const dim3 g((unsigned)grid_for(raw_w, 256), (unsigned)ncs);
            k_cols_seg<<<g, 256, 0, s>>>(offs, raw_w, raw_h, segc);
            k_cols_carry<<<grid_for(raw_w, 256), 256, 0, s>>>(raw_w, ncs, segc);
            k_cols_links<<<g, 256, 0, s>>>(offs, raw_w, raw_h, segc, pu);
        }
        // filled colour planes -> color_dev, then ten 5x5 median passes per channel
        // ping-ponging with scratch planes in the fill tables' memory (no longer needed); the
        // tenth pass writes each channel into its place in color_dev (interleaved for B, G, R)
        k_fill_resolve<<<grid_for(N, 256), 256, 0, s>>>(offs, raw_w, raw_h, nr, pr, nrf, prl, nd,
                                                        pu, colm, nch, color_dev);
        uint8_t* tmp = (uint8_t*)nr;  // N * 4 bytes >= nch planes
        const long long npairs = (long long)((raw_w + 1) >> 1) * raw_h;
        for (int ch = 0; ch < nch; ++ch) {
            uint8_t* a = color_dev + ch * N;
            uint8_t* b = tmp + ch * N;
            for (int it = 0; it < 9; ++it) {
                if (it & 1)
                    k_median5<<<grid_for(npairs, 256), 256, 0, s>>>(b, raw_w, raw_h, a, 1);
                else
                    k_median5<<<grid_for(npairs, 256), 256, 0, s>>>(a, raw_w, raw_h, b, 1);
            }
        }
        for (int ch = 0; ch < nch; ++ch)
            k_median5<<<grid_for(npairs, 256), 256, 0, s>>>(tmp + ch * N, raw_w, raw_h,
                                                            color_dev + ch, nch);
    }
    return check_launch("lfpm_virtual_depth_map");
}

int point_cloud_impl(const float* vd_dev, const uint8_t* color_dev, int nch, int raw_w,
                     int raw_h, float origin_x, float origin_y, float pixel_size,
                     float main_lens_f, float B, float bl0, uint8_t* records_dev,
                     size_t records_bytes, unsigned* count_dev, void* workspace_dev,
                     size_t workspace_bytes, void* stream);
}  // namespace
}  // namespace lfpm

extern "C" {

int lfpm_virtual_depth_map(const float* plane_dev, const float* centers_dev, const float* gray_dev,
                           int num_mis, int crop, float mla_base, int raw_w, int raw_h,
                           float* vd_dev, uint8_t* color_dev, float* range_dev,
                           void* workspace_dev, size_t workspace_bytes, void* stream) {
    return vd_map_impl(plane_dev, centers_dev, gray_dev, nullptr, num_mis, crop, mla_base, raw_w,
                       raw_h, vd_dev, color_dev, range_dev, workspace_dev, workspace_bytes,
                       stream);
}

int lfpm_virtual_depth_map_bgr(const float* plane_dev, const float* centers_dev,
                               const uint8_t* bgr_tiles_dev, int num_mis, int crop,
                               float mla_base, int raw_w, int raw_h, float* vd_dev,
                               uint8_t* color_dev, float* range_dev, void* workspace_dev,
                               size_t workspace_bytes, void* stream) {
    if (!bgr_tiles_dev)
        return set_error(LFPM_E_INVALID, "lfpm_virtual_depth_map_bgr: null colour crops");
    return vd_map_impl(plane_dev, centers_dev, nullptr, bgr_tiles_dev, num_mis, crop, mla_base,
                       raw_w, raw_h, vd_dev, color_dev, range_dev, workspace_dev,
                       workspace_bytes, stream);
}

int lfpm_vd_point_cloud(const float* vd_dev, const uint8_t* color_dev, int raw_w, int raw_h,
                        float origin_x, float origin_y, float pixel_size, float main_lens_f,
                        float B, float bl0, uint8_t* records_dev, size_t records_bytes,
                        unsigned* count_dev, void* workspace_dev, size_t workspace_bytes,
                        void* stream) {
    return point_cloud_impl(vd_dev, color_dev, 1, raw_w, raw_h, origin_x, origin_y, pixel_size,
                            main_lens_f, B, bl0, records_dev, records_bytes, count_dev,
                            workspace_dev, workspace_bytes, stream);
}

int lfpm_vd_point_cloud_bgr(const float* vd_dev, const uint8_t* color_bgr_dev, int raw_w,
                            int raw_h, float origin_x, float origin_y, float pixel_size,
                            float main_lens_f, float B, float bl0, uint8_t* records_dev,
                            size_t records_bytes, unsigned* count_dev, void* workspace_dev,
                            size_t workspace_bytes, void* stream) {
    if (!color_bgr_dev)
        return set_error(LFPM_E_INVALID, "lfpm_vd_point_cloud_bgr: null colour image");
    return point_cloud_impl(vd_dev, color_bgr_dev, 3, raw_w, raw_h, origin_x, origin_y,
                            pixel_size, main_lens_f, B, bl0, records_dev, records_bytes,
                            count_dev, workspace_dev, workspace_bytes, stream);
}

}  // extern "C"

namespace lfpm {
namespace {
int point_cloud_impl(const float* vd_dev, const uint8_t* color_dev, int nch, int raw_w,
                     int raw_h, float origin_x, float origin_y, float pixel_size,
                     float main_lens_f, float B, float bl0, uint8_t* records_dev,
                     size_t records_bytes, unsigned* count_dev, void* workspace_dev,
                     size_t workspace_bytes, void* stream) {
    if (raw_w <= 0 || raw_h <= 0 || (long long)raw_w * raw_h + 1 >= 0x7fffffffLL)
        return set_error(LFPM_E_INVALID, "lfpm_vd_point_cloud: bad frame %dx%d", raw_w, raw_h);
    if (!vd_dev || !records_dev || !count_dev || !workspace_dev)
        return set_error(LFPM_E_INVALID, "lfpm_vd_point_cloud: null pointer");
    const long long N = (long long)raw_w * raw_h;
    if (records_bytes < (size_t)N * 27)
        return set_error(LFPM_E_INVALID, "lfpm_vd_point_cloud: records %zu < %zu bytes",
                         records_bytes, (size_t)N * 27);
    const size_t nb = (size_t)((N + 1 + kScanChunk - 1) / kScanChunk);
    const size_t need = 2 * al((size_t)(N + 1) * 4) + al(nb * 4);
    if (workspace_bytes < need)
        return set_error(LFPM_E_INVALID, "lfpm_vd_point_cloud: workspace %zu < %zu bytes",
                         workspace_bytes, need);
    hipStream_t s = (hipStream_t)stream;
    unsigned char* ws = (unsigned char*)workspace_dev;
    int* flag = (int*)ws;
    int* offs = (int*)(ws + al((size_t)(N + 1) * 4));
    int* sums = (int*)(ws + 2 * al((size_t)(N + 1) * 4));
    k_pt_flag<<<grid_for(N + 1, 256), 256, 0, s>>>(vd_dev, N, flag);
    k_scan_sums<<<(int)nb, kScanThreads, 0, s>>>(flag, N + 1, sums);
    k_scan_sums_scan<<<1, kScanThreads, 0, s>>>(sums, (int)nb);
    k_scan_chunks<<<(int)nb, kScanThreads, 0, s>>>(flag, N + 1, sums, offs);
    k_pt_emit<<<grid_for(N, 256), 256, 0, s>>>(vd_dev, color_dev, nch, offs, raw_w, N, origin_x,
                                                origin_y, pixel_size, main_lens_f, B, bl0,
                                                records_dev, count_dev);
    return check_launch("lfpm_vd_point_cloud");
}
}  // namespace
}  // namespace lfpm

// Helpers shared with the composites of lfpm_fusion.hip (declared in lfpm_internal.h).
namespace lfpm {

size_t scan_sums_bytes(long long n) {
    return al((size_t)((n + kScanChunk - 1) / kScanChunk) * 4);
}

void exclusive_scan(const int* in, long long n, int* out, int* sums, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    const int nb = (int)((n + kScanChunk - 1) / kScanChunk);
    k_scan_sums<<<nb, kScanThreads, 0, s>>>(in, n, sums);
    k_scan_sums_scan<<<1, kScanThreads, 0, s>>>(sums, nb);
    k_scan_chunks<<<nb, kScanThreads, 0, s>>>(in, n, sums, out);
}

void median5_passes(const uint8_t* in, uint8_t* t0, uint8_t* t1, int raw_w, int raw_h, int passes,
                    uint8_t* out, int ostride, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    const long long npairs = (long long)((raw_w + 1) >> 1) * raw_h;
    const uint8_t* src = in;
    for (int it = 0; it + 1 < passes; ++it) {
        uint8_t* dst = (it & 1) ? t1 : t0;
        k_median5<<<grid_for(npairs, 256), 256, 0, s>>>(src, raw_w, raw_h, dst, 1);
        src = dst;
    }
    k_median5<<<grid_for(npairs, 256), 256, 0, s>>>(src, raw_w, raw_h, out, ostride);
}

}  // namespace lfpm
