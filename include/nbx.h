/*
 * nbx.h -- C ABI of libnbx, the MI355X (gfx950) Barnes-Hut step.
 *
 * The reference (tyreecepaul/n-body-simulation) has no plugin or FFI layer:
 * its boundary is the C++ class API of Simulation / CudaSimulation, the
 * `<n_bodies> <total_steps> <output_every> [output_path]` CLI and the NBS2
 * file. Each entry point below names the reference interface it replaces.
 *
 * Conventions
 *   - Every function returns NBX_OK (0) or an NBX_ERR_* code; the message of
 *     the last failure on the calling thread is nbx_last_error(). A wrapper
 *     that wants the reference's abort-on-error behaviour (CUDA_CHECK,
 *     src/cuda/DeviceBodies.cuh:18-31) prints it and exits(1).
 *   - A handle owns all device memory and one HIP stream; the caller owns
 *     host arrays, which are copied in and out. A handle is not thread-safe;
 *     separate handles are independent.
 *   - Arrays are structure-of-arrays, one entry per body: positions px/py,
 *     velocities vx/vy, accelerations ax/ay, mass m, radius r (float32) and
 *     the stable identity id (uint32) of src/Body.h:5-24.
 *   - The build reorders bodies every step (as QuadTree::build and
 *     CudaQuadTree::build do); "device order" is that order, ids identify
 *     bodies.
 */
#ifndef NBX_H
#define NBX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NBX_OK 0
#define NBX_ERR_ARG 1      /* bad argument (null pointer, n above capacity) */
#define NBX_ERR_HIP 2      /* a HIP runtime call failed */
#define NBX_ERR_CAPACITY 3 /* node, pair or sort-segment capacity exceeded */
#define NBX_ERR_STATE 4    /* call not valid in the handle's current state */

/* Force law: the reference's (fastInvSqrt, src/QuadTree.cpp:11-26) or an
 * accurate inverse square root. NBX_LAW_ACCURATE_QUAD (no reference
 * counterpart; compensated summation only) adds, for every far node, the
 * quadrupole term of the law m*d/(|d|(|d|^2+eps^2)) expanded about the
 * node's centre of mass, from the node's second mass moments -- a
 * closed-form far-field correction, so a coarser theta reaches a given
 * force error. */
#define NBX_LAW_REFERENCE 0
#define NBX_LAW_ACCURATE 1
#define NBX_LAW_ACCURATE_QUAD 2
/* NBX_LAW_LEARNED (no reference counterpart; compensated summation only):
 * the accurate law, where every far node of positive mass inside the
 * network's trained range adds its monopole plus the residual predicted by
 * the learned far-field network (nbx_set_mlp), and every far node outside
 * that range is opened instead, as a finer theta would open it. */
#define NBX_LAW_LEARNED 3

/* Per-body force summation: bit-exact replay of the reference's FP32 order
 * (src/QuadTree.cpp:264-389, four interleaved leaf sums), or FP32 with
 * compensated (Kahan) accumulation and a hardware reciprocal. */
#define NBX_SUM_REFERENCE 0
#define NBX_SUM_COMPENSATED 1

/* Collision resolution: the reference CPU's (src/Collision.cpp:135-349, pairs
 * one after another in pairLess order, each seeing the previous result) or
 * the reference GPU's (src/cuda/CudaCollision.cu, CollisionKernels.cuh:
 * every pair from the start-of-collide state, each body combining its own
 * side of its contacts; grid cell 2 x median radius x 1.01). Both bit-exact
 * against their reference. */
#define NBX_COLLIDE_SEQUENTIAL 0
#define NBX_COLLIDE_JACOBI 1

typedef struct nbx_sim nbx_sim;

/* Physics and execution parameters. Defaults (nbx_default_params) are the
 * reference's: dt 0.05, theta 1, epsilon 1, collisions on
 * (src/Simulation.cpp:7-9; src/cuda/CudaSimulation.cuh:39-43). */
typedef struct nbx_params {
  float dt;
  float theta;
  float epsilon;
  int32_t collisions; /* CudaSimulation::collisions */
  int32_t law;        /* NBX_LAW_* */
  int32_t summation;  /* NBX_SUM_* */
  int32_t use_graph;  /* replay captured HIP graphs in nbx_step (1) or launch
                         eagerly (0) */
  int32_t collide_mode; /* NBX_COLLIDE_* */
} nbx_params;

/* Message of the last failing call on this thread ("" if none). */
const char *nbx_last_error(void);

/* ABI version of this header (NBX_ABI_VERSION). */
#define NBX_ABI_VERSION 1
int nbx_abi_version(void);

/* Fills *p with the reference defaults. */
void nbx_default_params(nbx_params *p);

/* CudaSimulation(n) storage / CudaQuadTree(n) (src/cuda/CudaQuadTree.cu:
 * 269-334): a handle able to hold up to `capacity` bodies, with node storage
 * for 8*capacity+8 tree nodes. Allocates everything up front; steps never
 * allocate. */
int nbx_create(uint64_t capacity, nbx_sim **out);

/* nbx_create with creation flags (no reference counterpart: the reference
 * has one way to build its CudaSimulation). NBX_CREATE_NO_GRID_BARRIERS
 * replaces the kernels that wait on a grid barrier (the one-launch radix
 * sort with the bounding box and keys folded in, the collide's bucket scan
 * and component allocation) by their multi-launch forms, for handles that
 * share the GPU with concurrent work and cannot promise that all of a
 * launch's workgroups are resident at once. nbx_create(c, out) is
 * nbx_create_ex(c, flags, out) with this flag set when the environment
 * variable NBX_SORT is "multi". */
#define NBX_CREATE_NO_GRID_BARRIERS 1u
int nbx_create_ex(uint64_t capacity, uint32_t flags, nbx_sim **out);

/* Releases all device memory and the stream. Null is accepted. */
int nbx_destroy(nbx_sim *sim);

/* Simulation(n) / CudaSimulation(n) (src/Simulation.cpp:7-9,
 * src/cuda/CudaSimulation.cu:34-38): loads uniformDisc(n, seed)
 * (src/Utils.cpp:6-54), in generator order, and resets the frame counter. */
int nbx_init_disc(nbx_sim *sim, uint64_t n, uint32_t seed);

/* Initial conditions of the viewer's scenario menu, makeScenario(s, n, seed)
 * (src/viewer/Scenarios.cpp:29-62), bit for bit:
 *   NBX_SCENARIO_DISC          uniformDisc, as nbx_init_disc;
 *   NBX_SCENARIO_TWO_GALAXIES  TwoGalaxies: discs of n/2 (seed) and n - n/2
 *                              (seed + 1) bodies on a glancing course, the
 *                              second counter-rotating (Scenarios.cpp:36-49);
 *   NBX_SCENARIO_COLD_SQUARE   ColdSquare: unit bodies at rest, uniform in
 *                              [-3 sqrt(n), 3 sqrt(n)]^2 (Scenarios.cpp:50-57);
 * and one the reference does not have:
 *   NBX_SCENARIO_MERGER        two discs of n/2 and n - n/2 bodies (seeds
 *                              seed, seed + 1) rotating the same way, centres
 *                              4 R apart (R = 5 sqrt(n/2), the disc radius),
 *                              closing at 0.3 v_c(R) along a direction 15
 *                              degrees off the line of centres.
 * Ids are 0..n-1 in array order. The two-disc scenarios need n >= 2. */
#define NBX_SCENARIO_DISC 0
#define NBX_SCENARIO_TWO_GALAXIES 1
#define NBX_SCENARIO_COLD_SQUARE 2
#define NBX_SCENARIO_MERGER 3
int nbx_init_scenario(nbx_sim *sim, int scenario, uint64_t n, uint32_t seed);

/* CudaSimulation(const std::vector<Body>&) and CudaSimulation::reset
 * (src/cuda/CudaSimulation.cu:34-48), DeviceBodies::upload
 * (src/cuda/DeviceBodies.cu:25-50). ax/ay may be null (zero). */
int nbx_upload(nbx_sim *sim, uint64_t n, const float *px, const float *py,
               const float *vx, const float *vy, const float *ax,
               const float *ay, const float *m, const float *r,
               const uint32_t *id);

/* DeviceBodies::download (src/cuda/DeviceBodies.cu:52-72): state in device
 * order. Any pointer may be null to skip that field. */
int nbx_download(nbx_sim *sim, float *px, float *py, float *vx, float *vy,
                 float *ax, float *ay, float *m, float *r, uint32_t *id);

/* The output-frame path of sim_cuda (src/cuda/MainCuda.cu:90-105): state
 * scattered into id order (index k holds the body with id k; ids must be a
 * permutation of 0..n-1). Any pointer may be null. */
int nbx_download_by_id(nbx_sim *sim, float *px, float *py, float *vx,
                       float *vy, float *m, float *r);

/* Body count currently loaded. */
int nbx_count(nbx_sim *sim, uint64_t *n);

/* CudaSimulation's public dt/theta/epsilon/collisions members. */
int nbx_set_params(nbx_sim *sim, const nbx_params *p);
int nbx_get_params(nbx_sim *sim, nbx_params *p);

/* Simulation::step / CudaSimulation::step (src/Simulation.cpp:11-16,
 * src/cuda/CudaSimulation.cu:51-58), `steps` times: iterate -> collide (if
 * enabled) -> build -> forces. Returns after the work is queued and any
 * device-side error of these steps has been checked (one host sync). */
int nbx_step(nbx_sim *sim, uint32_t steps);

/* The individual phases (src/cuda/CudaSimulation.cu:60-74). */
int nbx_iterate(nbx_sim *sim); /* DeviceBodies.cuh:111-113 iterate() */
int nbx_collide(nbx_sim *sim); /* CudaCollider::collide */
int nbx_attract(nbx_sim *sim); /* build + forces */
int nbx_build(nbx_sim *sim);   /* CudaQuadTree::build (reorders bodies) */
int nbx_forces(nbx_sim *sim);  /* CudaQuadTree::computeForces */

/* Steps completed since the last init/upload (CudaSimulation::frame). */
int nbx_frame(nbx_sim *sim, uint64_t *frame);

/* Waits for queued work and reports any device-side error. */
int nbx_sync(nbx_sim *sim);

/* nbx_step without the final host sync: queues `steps` steps on the
 * handle's stream and returns; device-side errors of these steps surface
 * at the next nbx_sync / nbx_step. No reference counterpart (CudaSimulation
 * ::step always waits): for callers that keep several handles busy at once,
 * each on its own stream, e.g. an ensemble of independent simulations.
 * Handles stepped this way must be created with NBX_CREATE_NO_GRID_BARRIERS
 * (no kernel that waits on a grid barrier); any other handle is refused
 * with NBX_ERR_STATE. */
int nbx_step_async(nbx_sim *sim, uint32_t steps);

/* ---- Launch shape ----
 * The reference launches its force walk with one fixed block size
 * (forceKernel<<<blocksFor(n), kBlock>>>, src/cuda/CudaQuadTree.cu:511); the
 * best workgroup size of this walk depends on n and theta, so it is a
 * per-handle setting. */

/* Lanes per workgroup of the compensated force walk (the reference,
 * accurate and quadrupole laws): 64, 128 (the default) or 256. Every size
 * gives bit-identical forces; only speed differs. The reference-summation
 * walk, the counting walk and the learned law keep their own sizes. */
int nbx_set_walk_block(nbx_sim *sim, uint32_t lanes);
int nbx_get_walk_block(nbx_sim *sim, uint32_t *lanes);

/* Times `reps` (1..1000) force walks of the current tree with each
 * workgroup size, keeps the fastest (as nbx_set_walk_block would) and
 * returns it in *best; `ms` (may be null) receives each size's mean walk
 * time in milliseconds, in the order 64, 128, 256. With no tree built it
 * builds one first, which reorders the bodies as nbx_build does; the
 * accelerations are recomputed, bit-identical to the last force pass on
 * that tree. NBX_ERR_STATE under the reference summation or the learned
 * law, whose walks the setting does not reach. */
int nbx_tune_walk(nbx_sim *sim, uint32_t reps, uint32_t *best, float *ms);

/* ---- Inspection, for checks (CudaQuadTree::download / nodeCount /
 *      levelCount, src/cuda/CudaQuadTree.cuh:42-63). ---- */

/* Node count and level count of the last build. */
int nbx_tree_info(nbx_sim *sim, int64_t *nodes, int32_t *levels);

/* Node arrays of the last build (QuadTree.h:35-70 Node fields): centre of
 * mass, mass, quad centre/size, children, next, body range. Copies
 * min(nodes, cap) entries; any pointer may be null. */
int nbx_download_tree(nbx_sim *sim, int64_t cap, float *comX, float *comY,
                      float *mass, float *centerX, float *centerY,
                      float *size, int32_t *children, int32_t *next,
                      int32_t *start, int32_t *end);

/* Root quad {cx, cy, size} of the last build (Quad::newContaining). */
int nbx_root_quad(nbx_sim *sim, float quad[3]);

/* Keys of the last build in sorted order (hi = depths 0-15, lo = depths
 * 16-31, 2 bits per level) and the permutation it applied: perm[k] = the
 * pre-build device index of the body now at k. Any pointer may be null. */
int nbx_download_keys(nbx_sim *sim, uint32_t *hi, uint32_t *lo,
                      uint32_t *perm);

/* Last collide (Collision.cpp:135-349): grid cell size, number of pairs,
 * number of connected contact components and the largest one's pair count.
 * The cell is the broadphase's own choice (1.01 x the 5/8 quantile of a
 * sample of box widths, where the reference uses the exact median); the pair
 * set does not depend on it. After a Jacobi collide (NBX_COLLIDE_JACOBI):
 * the reference GPU's cell, the number of overlapping box pairs
 * (CudaCollider::lastAabbPairs), the number of bodies with a contact and the
 * largest contact count of one body. */
int nbx_collide_info(nbx_sim *sim, float *cell, int64_t *pairs,
                     int64_t *components, int64_t *largest);

/* The last collide's pairs (device indices, later/earlier as in
 * Collision.cpp:120-127) in the order they were resolved, grouped by
 * component: comp[k] is the component's smallest body index. Every body's
 * pairs appear in the reference's pairLess order (resolve() touches only its
 * two bodies, so any such order gives the sequential replay's result). */
int nbx_download_pairs(nbx_sim *sim, int64_t cap, int32_t *later,
                       int32_t *earlier, int32_t *comp, int64_t *count);

/* After a Jacobi collide (CudaCollider::lastContacts, CudaCollision.cu:
 * 309-312): per body, its number of contacts and of overlapping boxes in
 * which it is the later box, indexed as the bodies were at that collide.
 * NBX_ERR_STATE if the last collide ran in the sequential mode. */
int nbx_download_contacts(nbx_sim *sim, int32_t *contacts,
                          int32_t *aabb_as_later);

/* Runs the force walk once more in counting mode over the current tree and
 * returns far-node (mass > 0) and body-pair interaction totals, plus the
 * walk's wavefront work: node visits summed over waves and leaf-body terms
 * evaluated per wave (each costs all 64 lanes), so lane utilisation is
 * (far + pairs) / (64 * (visits + leaf terms)). Any pointer may be null.
 * Does not change the state; NBX_ERR_STATE before any tree build (or after
 * an upload). (No reference counterpart: instrumentation.) */
int nbx_count_interactions(nbx_sim *sim, uint64_t *far_nodes,
                           uint64_t *body_pairs, uint64_t *wave_visits,
                           uint64_t *wave_leaf_terms);

/* The compensated walk's work profile: runs the counting walk (as
 * nbx_count_interactions) and returns wave-level totals, each counted once
 * per wave whatever the number of lanes taking part:
 *   [0] sibling blocks popped        [1] lanes entering them (sum of masks)
 *   [2] node visits with a far lane  [3] far lanes at those visits
 *   [4] leaf visits, plain loop      [5] leaf records in plain loops
 *   [6] leaf visits, spread          [7] spread term evaluations
 *   [8] spread butterfly steps       [9] traversal stack pushes
 *  [10] empty nodes (no bodies) visited
 * Fills min(count, NBX_WALK_PROFILE_FIELDS) entries of out. NBX_ERR_STATE
 * under the reference summation or the learned law. (No reference
 * counterpart: instrumentation.) */
#define NBX_WALK_PROFILE_FIELDS 11
int nbx_walk_profile(nbx_sim *sim, uint64_t *out, uint32_t count);

/* ---- Timing (the per-phase event bench of src/cuda/CheckCuda.cu:677-712).
 * With profiling on, nbx_step / phase calls launch eagerly and record HIP
 * events on the handle's stream around each phase. ---- */
#define NBX_PHASE_ITERATE 0
#define NBX_PHASE_COLLIDE 1
#define NBX_PHASE_BUILD 2
#define NBX_PHASE_WALK 3
#define NBX_PHASES 4
int nbx_set_profiling(nbx_sim *sim, int on);
/* Accumulated milliseconds per phase and the number of steps they cover;
 * resets the accumulators. */
int nbx_phase_times(nbx_sim *sim, double ms[NBX_PHASES], uint64_t *steps);

/* The handle's HIP stream (a hipStream_t), for callers that order their own
 * work after it. */
int nbx_stream(nbx_sim *sim, void **stream);

/* ---- FP64 energy and momentum of the current state (no reference
 *      counterpart: the diagnostics of the precision study). Kinetic
 *      sum m|v|^2/2 and momentum sum m v are summed in FP64 on the host; the
 *      potential sum over pairs of m_i m_j phi(r_ij) is evaluated in FP64 on
 *      the GPU (O(N^2)) with phi(r) = -(pi/2 - atan(r/eps))/eps, the
 *      potential of the walk's law m*d/(|d|(|d|^2+eps^2)) (-1/r when
 *      eps2 = 0, coincident bodies then skipped). eps2 = epsilon^2. ---- */
int nbx_energy(nbx_sim *sim, double eps2, double *kinetic, double *potential,
               double momentum[2]);

/* ---- FP64 direct sum (CheckCuda.cu:456-475 exactTreeLawForces) on the
 *      GPU, the accuracy reference: accelerations of the listed target bodies
 *      (device indices) from all bodies, law m*d/(|d|(|d|^2+eps^2)). ---- */
int nbx_exact_forces(nbx_sim *sim, uint64_t n_targets,
                     const uint32_t *targets, double eps2, double *ax,
                     double *ay);

/* nbx_exact_forces without any device allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (harnesses that
 * forbid hipMalloc, tight serving or checking loops). Its partial sums,
 * results and target list live in one statically sized device array held by
 * each GPU until the process ends (about 8.5 MB), so at most
 * NBX_NOALLOC_MAX_TARGETS targets per call (more: NBX_ERR_ARG). Same kernels,
 * same results as nbx_exact_forces. Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (other threads, other handles'
 * streams) race and must be serialised by the caller. */
#define NBX_NOALLOC_MAX_TARGETS 8192
int nbx_exact_forces_noalloc(nbx_sim *sim, uint64_t n_targets,
                             const uint32_t *targets, double eps2, double *ax,
                             double *ay);

/* ---- Learned far-field correction (no reference counterpart). ----
 *
 * A network 10 -> 64 -> 64 -> 64 -> 2 with GELU (tanh form) after each
 * hidden layer, BF16 weights and activations with FP32 accumulation on the
 * matrix cores, FP32 master weights trained with Adam on the GPU. For a far
 * node (mass M > 0, centre of mass c, size s, second mass moments S about c)
 * and a target body at x it predicts the residual between the node's
 * monopole (the far term of QuadTree.cpp:276-296) and the exact force of
 * the node's bodies (the law of CheckCuda.cu:456-475). Inputs, with
 * d = c - x, r = |d|, w = 1/(r^2 + eps^2), q = S/M, k = 8 w:
 *   log2(M)/16, d.x sqrt(w), d.y sqrt(w), s/r, k q_xx, k q_xy, k q_yy,
 *   k d.q.d/r^2 (radial), k (d x q d)/r^2 (mixed), k tr(q) - radial;
 * output: the residual in units of M w mu / 2, mu = w tr(q) (monopole
 * magnitude times quadrupole scale). Trained range: s <= 1.2 r and the
 * target at least s/4 outside the node's square.
 *
 * Parameters (NBX_MLP_PARAMS floats, row-major [out][in]): W1 (64x10), b1,
 * W2 (64x64), b2, W3 (64x64), b3, W4 (2x64), b4. */
#define NBX_MLP_PARAMS 9154
typedef struct nbx_mlp nbx_mlp;

/* A network handle with room for `pair_capacity` training pairs (64 bytes
 * each) in device memory. */
int nbx_mlp_create(uint64_t pair_capacity, nbx_mlp **out);
int nbx_mlp_destroy(nbx_mlp *mlp);

/* Master parameters in and out; setting them resets the Adam state. */
int nbx_mlp_set_params(nbx_mlp *mlp, const float *params);
int nbx_mlp_get_params(nbx_mlp *mlp, float *params);

/* Training set: number of pairs held, emptied by nbx_mlp_clear. */
int nbx_mlp_count(nbx_mlp *mlp, uint64_t *count);
int nbx_mlp_clear(nbx_mlp *mlp);

/* Appends the training pairs of the simulation's last tree (built with
 * node moments: law NBX_LAW_ACCURATE_QUAD or NBX_LAW_LEARNED): for each
 * listed target body (device index), every node of positive mass that a
 * walk at `theta` accepts inside the trained range (nodes outside it are
 * opened); inputs as the walk computes them, the target residual in FP64
 * over the node's bodies, and a loss weight: the squared ratio of the
 * output unit M w mu to the magnitude of the sum of the target's recorded
 * monopoles, so the loss measures each target's relative force error. *added (may be null) gets the number appended;
 * NBX_ERR_CAPACITY if the set overflowed (the excess is dropped). At most
 * 2^22 targets per call. */
int nbx_mlp_generate(nbx_mlp *mlp, nbx_sim *sim, float theta,
                     uint64_t n_targets, const uint32_t *targets,
                     uint64_t *added);

/* The (target body, node) pairs the last nbx_mlp_generate appended, in the
 * set's order (device indices and node indices of that tree): up to cap
 * written, *count = how many it appended. For checks. */
int nbx_mlp_last_pairs(nbx_mlp *mlp, uint64_t cap, uint32_t *target,
                       int32_t *node, uint64_t *count);

/* Replaces the training set with host pairs (inputs count x 10, targets
 * count x 2, loss weights count; null weights: all 1), or reads pairs
 * [first, first + count) back (weights may be null). */
int nbx_mlp_upload(nbx_mlp *mlp, uint64_t count, const float *features,
                   const float *labels, const float *weights);
int nbx_mlp_download(nbx_mlp *mlp, uint64_t first, uint64_t count,
                     float *features, float *labels, float *weights);

/* Loss -- the weighted mean squared error sum_k w_k |y_k - t_k|^2 /
 * (count mean(w)) -- and its gradient (NBX_MLP_PARAMS floats) over pairs
 * [first, first + count) of the set, without changing the parameters: the
 * training kernels' backward pass, for checks. */
int nbx_mlp_gradient(nbx_mlp *mlp, uint64_t first, uint64_t count,
                     float *grad, float *loss);

/* `steps` Adam steps (beta 0.9 / 0.999, eps 1e-8, learning rate lr) on the
 * weighted loss of batches of `batch` pairs (weights normalised by the
 * whole set's mean) taken in the order of a seeded permutation of the set;
 * losses (may be null) gets each step's batch loss. */
int nbx_mlp_train(nbx_mlp *mlp, uint32_t steps, uint32_t batch, float lr,
                  uint64_t seed, float *losses);

/* Network outputs (count x 2) for host inputs (count x 10). */
int nbx_mlp_predict(nbx_mlp *mlp, uint64_t count, const float *features,
                    float *out);

/* Copies the network's current weights into the simulation, for
 * NBX_LAW_LEARNED (later changes to the network need another call). */
int nbx_set_mlp(nbx_sim *sim, nbx_mlp *mlp);

/* ---- Ensembles: many independent simulations stepped together (no
 * reference counterpart; each instance is a CudaSimulation of its own,
 * src/cuda/CudaSimulation.cuh:25-51, stepped as CudaSimulation::step,
 * CudaSimulation.cu:51-58). A step of every instance is two dispatches
 * whatever the number of instances: a persistent kernel in which each
 * workgroup takes whole instances (largest first) through iterate, collide
 * and build, and one walk over all instances' bodies; steps replay a
 * captured HIP graph. Results are bit-identical to stepping each instance
 * with an nbx_sim of its own (reference or accurate law, compensated
 * summation, sequential collide). ---- */
typedef struct nbx_ensemble nbx_ensemble;

/* `count` instances (at most 65536) with room for capacity[k] <= 2^20
 * bodies each; all memory allocated here. Instances start empty. */
int nbx_ensemble_create(uint32_t count, const uint64_t *capacity,
                        nbx_ensemble **out);
int nbx_ensemble_destroy(nbx_ensemble *ens);

/* Loads instance k: makeScenario's initial condition (NBX_SCENARIO_*, as
 * nbx_init_scenario) or given bodies (as nbx_upload; ax/ay may be null). */
int nbx_ensemble_init(nbx_ensemble *ens, uint32_t k, int scenario, uint64_t n,
                      uint32_t seed);
int nbx_ensemble_upload(nbx_ensemble *ens, uint32_t k, uint64_t n,
                        const float *px, const float *py, const float *vx,
                        const float *vy, const float *ax, const float *ay,
                        const float *m, const float *r, const uint32_t *id);

/* Instance k's bodies in device order (any pointer may be null). */
int nbx_ensemble_download(nbx_ensemble *ens, uint32_t k, float *px, float *py,
                          float *vx, float *vy, float *ax, float *ay, float *m,
                          float *r, uint32_t *id);

/* Parameters shared by all instances: law NBX_LAW_REFERENCE or
 * NBX_LAW_ACCURATE, compensated summation, the sequential collide. */
int nbx_ensemble_set_params(nbx_ensemble *ens, const nbx_params *p);

/* `steps` steps of every instance; returns after they finished, with
 * NBX_ERR_CAPACITY if an instance overflowed its node or pair storage. */
int nbx_ensemble_step(nbx_ensemble *ens, uint32_t steps);

/* Launch shape of nbx_ensemble_step: `walk_lanes` lanes per workgroup of
 * the walk (64, 128 or 256; the default 128) and `prepare_groups`
 * workgroups of the per-instance kernel (1..count; the default is twice
 * the device's compute units, at most count). 0 keeps a setting as it is.
 * Results are bit-identical for every choice; the tuner times them. */
int nbx_ensemble_set_launch(nbx_ensemble *ens, uint32_t walk_lanes,
                            uint32_t prepare_groups);
int nbx_ensemble_get_launch(nbx_ensemble *ens, uint32_t *walk_lanes,
                            uint32_t *prepare_groups);

#ifdef __cplusplus
}
#endif

#endif
