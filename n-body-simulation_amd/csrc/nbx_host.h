// nbx_host.h -- host-only helpers of libnbx (compiled by g++, not hipcc).
#pragma once

#include <cstdint>

namespace nbx_host {

// uniformDisc(n, seed) (src/Utils.cpp:6-54): one heavy central body, the rest
// uniform by area in the annulus [25, 5*sqrt(n)] on circular orbits, sorted by
// |x|^2; ids assigned before the sort. Arrays of length n, generator order.
// Built with -O2 -ffp-contract=off against the same libstdc++ <random> and
// glibc sinf/cosf as the reference, so the bits match its CPU build.
void uniformDisc(uint64_t n, uint32_t seed, float *px, float *py, float *vx,
                 float *vy, float *m, float *r, uint32_t *id);

} // namespace nbx_host
