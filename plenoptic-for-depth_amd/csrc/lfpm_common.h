// lfpm_common.h — shared host/device definitions of the MI355X micro-image PatchMatch solver.
#pragma once
#include <stdint.h>

#include <cmath>

#if defined(__HIPCC__)
#define LFPM_HD __host__ __device__ __forceinline__
#else
#define LFPM_HD inline
#endif

namespace lfpm {

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers: As Easy as 1, 2, 3",
// SC'11).  Replaces the clock-seeded curand XORWOW of the reference
// (MIStereo/AdaptMIPMFrame.cu:501), so a solve is a pure function of (seed, frame_uid).
LFPM_HD void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                           uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Draw-slot contract of the solver stream.  For pixel uid P (v*crop*crop + y*crop + x) and step S
// (0 = random init, 1 + 2*iter + phase = update), slot s is word s%4 of
// Philox(ctr = (P, S, s/4, 0), key = (seed, frame_uid)).
enum : uint32_t {
    SLOT_INIT_DEPTH = 0,
    SLOT_INIT_MARSAGLIA = 4,   // attempt a: slots 4+2a, 5+2a
    SLOT_VIEW_DRAW = 0,        // 15 view-sampling draws: slots 0..14
    SLOT_REF_DEPTH_RAND = 16,
    SLOT_REF_DEPTH_PERT = 17,
    SLOT_REF_ANGLE = 20,       // 20, 21, 22
    SLOT_REF_MARSAGLIA = 24,   // attempt a: slots 24+2a, 25+2a
    SLOT_WARM_JITTER = 40,     // ACMM warm-seed disparity jitter (init step)
    SLOT_WARM_ANGLE = 44,      // ACMM warm-seed normal perturbation, 44..46 (init step)
    MARSAGLIA_ATTEMPTS = 16,
    VIEW_DRAWS = 15,
};

// ((x >> 8) + 0.5) * 2^-24 — exact in FP32, in (0, 1).
LFPM_HD float u01(uint32_t x) { return ((float)(x >> 8) + 0.5f) * 5.9604644775390625e-08f; }

// Hex lattice centre of lattice cell (col i, row j); odd rows are shifted by half a pitch.
// The ideal centre is snapped to the nearest point whose crop origin c - (crop-1)/2 is an integer,
// so the integer crop of the reference (DepthSolver.cpp:1894-1903, 1980-1983) is centred exactly
// on the lens and local MI coordinates differ by exactly c_ref - c_nei.  Computed in double and
// rounded once so host builders agree bit for bit.
inline void lattice_center(float pitch, int crop, int i, int j, float& cx, float& cy) {
    const double p = (double)pitch;
    const double h = (crop - 1) * 0.5;
    const double ix = p * ((double)i + 0.5 * (double)(1 + (j & 1)));
    const double iy = p * (0.5 + 0.8660254037844386 * (double)j);
    cx = (float)(std::floor(ix - h + 0.5) + h);
    cy = (float)(std::floor(iy - h + 0.5) + h);
}

} // namespace lfpm
