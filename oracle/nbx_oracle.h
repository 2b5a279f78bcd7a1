/*
 * nbx_oracle.h -- CPU oracle for the Barnes-Hut step (TEST INFRASTRUCTURE).
 *
 * This is a plain restatement of the reference CPU algorithm
 * (tyreecepaul/n-body-simulation, src/ C++ files), used only as the checker:
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.
 * The product (the HIP library) never links, loads or calls it.
 *
 * All arrays are structure-of-arrays, caller-owned, length n unless stated.
 * Every function is single-threaded and deterministic. Build flags are fixed
 * by oracle/Makefile (-O2 -ffp-contract=off, baseline x86-64: no FMA), which
 * is the configuration whose bits the HIP kernels reproduce.
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   or_uniform_disc    reference build (oracle/_ref) + golden frame-0 fixture
 *   or_scenario        TwoGalaxies / ColdSquare: reference build (makeScenario)
 *                      + golden frame-0 fixtures; the merger, which the
 *                      reference does not have: closed forms (centres 4 R
 *                      apart, bulk velocities, zero total momentum for odd
 *                      and even n) and its two discs = or_uniform_disc
 *   or_iterate         closed form + reference build
 *   or_root_quad       exact min/max + reference tree root
 *   or_keys            reference QuadTree node ranges (every leaf), numpy
 *                      restatement of the descent
 *   or_sort_perm       stable-sort semantics (numpy lexsort)
 *   or_build_tree      reference QuadTree::build node arrays, bit-exact, on
 *                      key-presorted input (no reordering by the reference)
 *   or_accelerations   reference QuadTree::acceleration, bit-exact; theta->0
 *                      limit vs the FP64 direct sum; fastInvSqrt exhaustive
 *   or_collide         reference collide() output bit-exact; pair set vs an
 *                      O(N^2) brute-force sweep in numpy
 *   or_collide_jacobi  reference GPU collide semantics: bit-exact against a
 *                      host build of the reference's CollisionKernels.cuh
 *                      (oracle/_ref) driven in CudaCollision.cu's order
 *   or_step            reference Simulation::step (per phase), invariants
 *   or_exact_forces    closed-form two/three-body cases
 *   or_mlp_*           see nbx_oracle_mlp.cpp (finite differences, closed
 *                      forms, reduction to the plain walk)
 */
#ifndef NBX_ORACLE_H
#define NBX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Force law / summation selectors shared with the HIP library. */
enum { OR_LAW_REFERENCE = 0, OR_LAW_ACCURATE = 1 };

/* Utils.cpp:6-54 uniformDisc(n, seed): bodies in generator order (sorted by
 * |x|^2), ids assigned before the sort. */
void or_uniform_disc(uint64_t n, uint32_t seed, float *px, float *py,
                     float *vx, float *vy, float *m, float *r, uint32_t *id);

/* src/viewer/Scenarios.cpp:29-62 makeScenario(kind, n, seed), ids 0..n-1:
 * kind 0 Disc, 1 TwoGalaxies, 2 ColdSquare; kind 3 is the precision-study
 * merger (two same-sense discs of n/2 and n - n/2 bodies, seeds seed and
 * seed + 1, centres at (-2R, 0) and (2R, 0), R = 5 sqrt(n/2), bulk velocities
 * +/-0.15 v_c (cos 15deg, sin 15deg), v_c = sqrt((1e6 + n/2 - 1) / R)).
 * Returns 0, or -1 for an unknown kind or a two-disc kind with n < 2. */
int or_scenario(int kind, uint64_t n, uint32_t seed, float *px, float *py,
                float *vx, float *vy, float *m, float *r, uint32_t *id);

/* Body.h:20-23 Body::update / Simulation.cpp:44-50: v += a*dt; x += v*dt. */
void or_iterate(uint64_t n, float dt, float *px, float *py, float *vx,
                float *vy, const float *ax, const float *ay);

/* QuadTree.cpp:30-51 Quad::newContaining -> quad = {cx, cy, size}. */
void or_root_quad(uint64_t n, const float *px, const float *py, float quad[3]);

/* Float-descent quadrant key (QuadTree.h:14-25 intoQuadrant + the partition
 * predicates of QuadTree.cpp:60-78): 2 bits per level, digit = 2*(y>=cy) +
 * (x>=cx); depth 0 digit in bits 31:30 of hi, depths 16..31 in lo. */
void or_keys(uint64_t n, const float *px, const float *py,
             const float quad[3], uint32_t *hi, uint32_t *lo);

/* Stable sort of indices by (hi, lo). */
void or_sort_perm(uint64_t n, const uint32_t *hi, const uint32_t *lo,
                  uint32_t *perm);

/* QuadTree.cpp:154-262 build + accumulateLeaf + propagate, on bodies already
 * in (hi, lo) order. Node arrays need `capacity` entries. Returns the node
 * count, or -1 if capacity would be exceeded. levels_out gets the number of
 * levels. Nodes at depth 32 stay leaves (keys hold 32 levels). */
int64_t or_build_tree(uint64_t n, const float *px, const float *py,
                      const float *m, const uint32_t *hi, const uint32_t *lo,
                      const float quad[3], int64_t capacity, float *comX,
                      float *comY, float *mass, float *centerX,
                      float *centerY, float *size, int32_t *children,
                      int32_t *next, int32_t *start, int32_t *end,
                      int32_t *levels_out);

/* QuadTree.cpp:264-389 acceleration() for every body (positions px,py),
 * tree from or_build_tree. law = OR_LAW_REFERENCE uses fastInvSqrt
 * (QuadTree.cpp:11-26); OR_LAW_ACCURATE uses 1/sqrtf. Also returns per-body
 * interaction counts (far nodes, body pairs) when the pointers are non-null. */
void or_accelerations(uint64_t n, const float *px, const float *py,
                      const float *m, int64_t nodes, const float *comX,
                      const float *comY, const float *mass,
                      const float *size, const int32_t *children,
                      const int32_t *next, const int32_t *start,
                      const int32_t *end, float theta, float epsilon,
                      int law, float *ax, float *ay, uint32_t *far_count,
                      uint32_t *pair_count);

/* Collision.cpp:135-349 collide(): detects the pair list, orders it by
 * pairLess and resolves it sequentially (Collision.cpp:16-84). pos/vel are
 * updated in place. The ordered pairs are written to later[]/earlier[] (up to
 * pair_cap); returns the pair count (which may exceed pair_cap). cell_out
 * (may be null) receives the grid cell size. */
int64_t or_collide(uint64_t n, float *px, float *py, float *vx, float *vy,
                   const float *m, const float *r, int32_t *later,
                   int32_t *earlier, int64_t pair_cap, float *cell_out);

/* The reference GPU collide (src/cuda/CudaCollision.cu:205-296 with
 * src/cuda/CollisionKernels.cuh:104-289): every pair resolved from the
 * start-of-collide state, each body combining its own side of its contacts
 * (one contact: that pair's result; several: start + sum of changes), in the
 * kernels' scan and summation order: small boxes over the 3x3 stencil of a
 * row-major 64-bit cell key sorted stably with index payload, then every
 * large box; each large box against chunks of 4096 bodies scanned by 128
 * strided lanes, a fixed tree reduction per chunk, chunks summed in order.
 * `cell` is the grid cell (the reference: 2 x median radius x 1.01, fixed at
 * construction). contacts / aabb_as_later (may be null) receive per body the
 * number of contacts and of overlapping boxes where it is the later box.
 * Returns the number of large boxes. */
int64_t or_collide_jacobi(uint64_t n, float cell, float *px, float *py,
                          float *vx, float *vy, const float *m, const float *r,
                          int32_t *contacts, int32_t *aabb_as_later);

/* CudaCollider's cell (CudaCollision.cu:197-209): 2 x the median radius
 * (nth_element at n/2) x 1.01; 1 when that is not positive. */
float or_jacobi_cell(uint64_t n, const float *r);

/* Simulation.cpp:11-16 step(): iterate -> collide (if enabled) -> build ->
 * accelerations. Bodies are permuted into key order by the build (all nine
 * arrays). Returns 0, or -1 on node-capacity overflow. */
int or_step(uint64_t n, float dt, float theta, float epsilon, int collisions,
            int law, float *px, float *py, float *vx, float *vy, float *ax,
            float *ay, float *m, float *r, uint32_t *id);

/* FP64 direct sum with the tree's force law m*d/(|d|(|d|^2+eps^2))
 * (CheckCuda.cu:456-475) for the listed target bodies; self term skipped. */
void or_exact_forces(uint64_t n, const float *px, const float *py,
                     const float *m, uint64_t n_targets,
                     const uint32_t *targets, double eps2, double *ax,
                     double *ay);

/* ---- Learned far-field correction (nbx_oracle_mlp.cpp). No reference
 * counterpart; the term it corrects is the far-node monopole of
 * QuadTree.cpp:276-296 and its judge the exact sum of CheckCuda.cu:456-475.
 *
 * Network: 10 -> 64 -> 64 -> 64 -> 2, GELU (tanh form) after each hidden
 * layer. Parameters, flat, row-major [out][in]: W1 (64x10), b1, W2 (64x64),
 * b2, W3 (64x64), b3, W4 (2x64), b4 -- OR_MLP_PARAMS floats.
 *
 * A pair is (target body t, accepted node k with mass M > 0, centre of mass
 * c, square centre C and size s, second mass moments S about c). With
 * d = c - t, r^2 = |d|^2, w = 1/(r^2 + eps^2), q = S / M and
 * k = OR_MLP_MOMENT_SCALE * w, the inputs are
 *   f0 = log2(M) / 16, f1 = d.x sqrt(w), f2 = d.y sqrt(w), f3 = s / r,
 *   f4..f6 = k (qxx, qxy, qyy),
 *   f7 = k d^T q d / r^2, f8 = k (d x q d) / r^2, f9 = k tr q - f7
 * (q in the frame of d: radial, mixed, tangential), and the target is the
 * residual R = sum over the node's bodies of the law minus the node's
 * monopole in units of u = M w mu, mu = w tr(q) (the monopole's magnitude
 * times the quadrupole scale): R / u * OR_MLP_OUT_SCALE (0 when u = 0, a
 * node whose mass sits at its centre of mass). The trained
 * range: s^2 <= OR_MLP_MAX_RATIO^2 r^2 and the target at least
 * OR_MLP_BOX_MARGIN * s from the node's square (FP32, as the walk tests it);
 * far nodes outside it are opened instead. */
#define OR_MLP_IN 10
#define OR_MLP_HIDDEN 64
#define OR_MLP_OUT 2
#define OR_MLP_PARAMS 9154
#define OR_MLP_MAX_RATIO 1.2
#define OR_MLP_BOX_MARGIN 0.25
#define OR_MLP_MOMENT_SCALE 8.0
#define OR_MLP_OUT_SCALE 0.5

/* Stackless preorder walk (QuadTree.cpp:264-389's order) at `theta` for
 * each target, recording every accepted node of positive mass; with
 * use_range, far nodes outside the trained range are opened instead (and
 * not recorded). Writes up to cap pairs; returns the pair count. */
int64_t or_mlp_pairs(uint64_t n, const float *px, const float *py,
                     int64_t nodes, const float *comX, const float *comY,
                     const float *mass, const float *centerX,
                     const float *centerY, const float *size,
                     const int32_t *children, const int32_t *next,
                     uint64_t n_targets, const uint32_t *targets, float theta,
                     int use_range, int64_t cap, uint32_t *out_target,
                     int32_t *out_node);

/* Inputs (count x 10) and regression targets (count x 2) of the pairs,
 * moments and residuals in FP64 from the node's bodies [start, end), and
 * (weight may be null) each pair's loss weight (u / |A|)^2, u the output
 * unit above and A the sum of the monopoles of the listed pairs with the
 * same target: an output error e costs e u / |A| of that target's force. */
void or_mlp_features(const float *px, const float *py, const float *m,
                     const float *comX, const float *comY, const float *mass,
                     const float *size, const int32_t *start,
                     const int32_t *end, uint64_t count,
                     const uint32_t *targets, const int32_t *nodes,
                     double eps2, float *feat, float *label, float *weight);

/* Network outputs (count x 2). bf16 != 0 rounds weights, inputs and hidden
 * activations to BF16 (round to nearest even) and computes in FP32; 0
 * computes exactly in double. */
void or_mlp_forward(const float *params, uint64_t count, const float *feat,
                    int bf16, float *out);

/* Weighted mean squared error, sum_k w_k |y_k - t_k|^2 / (count wnorm)
 * (weight null: every w_k = 1), and its gradient with respect to every
 * parameter (OR_MLP_PARAMS doubles). bf16 as above, with the backward
 * operands (output and hidden deltas) rounded to BF16 too. */
void or_mlp_loss_grad(const float *params, uint64_t count, const float *feat,
                      const float *label, const float *weight, double wnorm,
                      int bf16, double *loss, double *grad);

/* The learned-law walk for the listed targets, in FP64: accurate law; far
 * nodes inside the trained range add their monopole plus the network's
 * residual (params may be null: monopole only), far nodes outside it are
 * opened; leaves add their bodies. Per-target interaction counts optional. */
void or_mlp_accelerations(uint64_t n, const float *px, const float *py,
                          const float *m, int64_t nodes, const float *comX,
                          const float *comY, const float *mass,
                          const float *centerX, const float *centerY,
                          const float *size, const int32_t *children,
                          const int32_t *next, const int32_t *start,
                          const int32_t *end, const float *params, int bf16,
                          uint64_t n_targets, const uint32_t *targets,
                          float theta, double eps2, double *ax, double *ay,
                          uint32_t *far_count, uint32_t *pair_count);

/* Round to the nearest BF16 value (ties to even), as the GPU converts. */
float or_bf16_round(float x);

/* QuadTree.cpp:11-26 fastInvSqrt, exposed for the exhaustive known-answer
 * test: out[i] = fisr(bits-reinterpreted in[i]) for i in [0, count). */
void or_fast_inv_sqrt_bits(uint64_t count, const uint32_t *in, uint32_t *out);

#ifdef __cplusplus
}
#endif

#endif
