// nbx_host.h -- host-only helpers of libnbx (compiled by g++, not hipcc).
#pragma once

#include <cstdint>

namespace nbx_host {

// Initial conditions. Each fills arrays of length n (positions, velocities,
// masses, radii, ids) in generator order. They are built with -O2
// -ffp-contract=off against the same libstdc++ <random> and glibc
// sinf/cosf/sqrtf/cbrtf as the reference, so the bits match its CPU build.

// uniformDisc(n, seed) (src/Utils.cpp:6-54): one heavy central body, the rest
// uniform by area in the annulus [25, 5*sqrt(n)] on circular orbits, sorted by
// |x|^2; ids assigned before the sort.
void uniformDisc(uint64_t n, uint32_t seed, float *px, float *py, float *vx,
                 float *vy, float *m, float *r, uint32_t *id);

// makeScenario(TwoGalaxies, n, seed) (src/viewer/Scenarios.cpp:36-49): discs
// of n/2 (seed) and n - n/2 (seed + 1) bodies at (-sep/2, -R/3) and
// (sep/2, R/3), sep = 2.5 R, R = 5 sqrt(n/2), closing at 2v with
// v = 0.3 sqrt(M/sep), the second disc counter-rotating; ids 0..n-1.
void twoGalaxies(uint64_t n, uint32_t seed, float *px, float *py, float *vx,
                 float *vy, float *m, float *r, uint32_t *id);

// makeScenario(ColdSquare, n, seed) (src/viewer/Scenarios.cpp:50-57): unit
// bodies at rest, uniform in [-3 sqrt(n), 3 sqrt(n)]^2; ids 0..n-1. Each
// position draws y before x: the reference writes Vec2(coord(rng),
// coord(rng)), and its g++ build evaluates the second argument first.
void coldSquare(uint64_t n, uint32_t seed, float *px, float *py, float *vx,
                float *vy, float *m, float *r, uint32_t *id);

// The precision-study merger (not in the reference): uniformDisc(n/2, seed)
// and uniformDisc(n - n/2, seed + 1), both rotating the same way, centred at
// (-2 R, 0) and (2 R, 0) with R = 5 sqrt(n/2) (centres 4 R apart), closing
// at 0.3 v_c along a direction 15 degrees off the line of centres
// (+/- 0.15 v_c (cos 15, sin 15) each), v_c = sqrt(M / R) the circular speed
// at R of a disc of mass M = 1e6 + n/2 - 1; ids 0..n-1.
void galaxyMerger(uint64_t n, uint32_t seed, float *px, float *py, float *vx,
                  float *vy, float *m, float *r, uint32_t *id);

} // namespace nbx_host
