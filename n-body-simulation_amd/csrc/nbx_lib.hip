// nbx_lib.hip -- libnbx runtime: device buffers, launch sequences, HIP graphs
// and the C ABI declared in include/nbx.h.
//
// One handle = one simulation resident in HBM: body arrays (double-buffered
// so the build's permutation is a single gather pass), collision scratch,
// sort scratch and tree storage, all allocated once at nbx_create. A step
// is a fixed sequence of kernel launches on the handle's stream whose sizes
// live on the device (Ctl), so it needs no host round trip and can be
// captured into a HIP graph (one per buffer parity).
#include "nbx.h"

#include "k_collide.h"
#include "k_ensemble.h"
#include "k_jacobi.h"
#include "k_sort.h"
#include "k_tree.h"
#include "k_walk.h"
#include "nbx_device.h"
#include "nbx_host.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace nbx;

// Device memory of nbx_exact_forces_noalloc, whose callers cannot allocate:
// per-(body chunk, target) FP64 partial sums, the per-target results and the
// target list, sized for kNoAllocTargets targets. Each GPU holds its own
// copy until the process ends; in host code the name is only a stand-in, so
// its per-device address (resolved once per device with hipGetSymbolAddress)
// is what the host passes to kernels and copies.
constexpr uint32_t kNoAllocTargets = 8192;
constexpr size_t kScratchPartials = static_cast<size_t>(kExactChunks) * kNoAllocTargets;
constexpr size_t kScratchEntries = kScratchPartials + kNoAllocTargets + kNoAllocTargets / 4;
static __device__ __align__(256) double2 scratch[kScratchEntries];

namespace {

thread_local std::string g_error;

int setError(int code, const std::string &msg) {
  g_error = msg;
  return code;
}

#define NBX_HIP(call)                                                         \
  do {                                                                        \
    hipError_t nbxErr_ = (call);                                              \
    if (nbxErr_ != hipSuccess) {                                              \
      return setError(NBX_ERR_HIP, std::string(#call) + " failed: " +         \
                                       hipGetErrorString(nbxErr_));           \
    }                                                                         \
  } while (0)

#define NBX_TRY(expr)                                                         \
  do {                                                                        \
    int nbxRc_ = (expr);                                                      \
    if (nbxRc_ != NBX_OK) {                                                   \
      return nbxRc_;                                                          \
    }                                                                         \
  } while (0)

uint32_t blocksFor(uint64_t n, uint64_t per = kBlock) {
  return static_cast<uint32_t>(std::max<uint64_t>(1, (n + per - 1) / per));
}

uint64_t pow2AtLeast(uint64_t v) {
  uint64_t p = 1;
  while (p < v) {
    p <<= 1;
  }
  return p;
}

// Per-device address of `scratch`, resolved with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
double2 *g_scratchPtr[kMaxDevices] = {};

int resolveScratch(int dev, double2 **out) {
  if (dev < 0 || dev >= kMaxDevices) {
    return setError(NBX_ERR_ARG, "device index out of range");
  }
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (!g_scratchPtr[dev]) {
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
      return setError(NBX_ERR_HIP, "hipGetSymbolAddress(scratch) failed");
    }
    g_scratchPtr[dev] = static_cast<double2 *>(ptr);
  }
  *out = g_scratchPtr[dev];
  return NBX_OK;
}

} // namespace

struct nbx_sim {
  int device = 0; // the HIP device current at nbx_create
  uint64_t cap = 0;
  uint64_t n = 0;
  int cur = 0; // body buffer holding the state
  uint64_t frame = 0;
  nbx_params params{};
  double massAbsSum = 0; // sum of |m| of the bodies held (masses never change)
  bool massPositive = false; // every body mass in (0, FLT_MAX]
  float jacobiCell = 1.0f; // reference GPU collide cell (radii never change)
  float gridCell[2] = {1.0f, 1.0f}; // sequential collide cell and inverse
  JacAcc *jacPartials = nullptr;
  uint32_t jacMaxLarge = 0;
  bool jacobiCollided = false; // last collide ran in the Jacobi mode
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;

  // bodies
  float *px[2]{}, *py[2]{}, *vx[2]{}, *vy[2]{}, *m[2]{}, *r[2]{};
  uint32_t *id[2]{};
  float *ax = nullptr, *ay = nullptr;

  // collide
  uint64_t buckets = 0; // power of two >= 2 * capacity
  uint32_t *bucketOf = nullptr, *bucketRank = nullptr, *bucketCount = nullptr,
           *bucketStart = nullptr;
  uint32_t *scanTiles = nullptr;
  float4 *grid = nullptr; // per entry: box, then {cx, cy, index} (32 B)
  int32_t *largeList = nullptr, *ufParent = nullptr;
  float4 *largeBox = nullptr; // {minX, maxX, minY, maxY} per large body
  uint32_t pairCap = 0;
  int4 *pairs = nullptr, *segPairs = nullptr;
  int32_t *pairComp = nullptr;
  uint32_t *compCount = nullptr, *compStart = nullptr;
  int4 *compInfo = nullptr;
  int32_t *bigList = nullptr;
  int4 *tierList = nullptr; // kTierQueues x kWaveTiers queues of tierCap {start, size, component}
  uint32_t tierCap = 0;
  int32_t *bodySlot = nullptr; // per body: slot in a big component's replay, -1 at rest

  // tree
  uint32_t *keyHi[2]{}, *vals[2]{};
  uint32_t *keyLo = nullptr, *sortHist = nullptr, *sortRowTotal = nullptr;
  int keySource = 0;      // buffer whose bodies the last build keyed
  bool keyLowFilled = false; // every low key word of the last build stored
  uint32_t *sortBarrier = nullptr; // grid barrier words of k_sort_onelaunch
  uint32_t sortGrid = 0; // its maximum workgroups; 0: three launches a pass
  uint32_t barrierSpins = kBarrierSpins; // poll bound of the grid barriers
  uint32_t *tileCount = nullptr;
  int64_t nodeCap = 0;
  int32_t *splitList = nullptr, *nodeOfSplit = nullptr;
  float4 *nodeF = nullptr;
  float4 *nodeQ = nullptr; // second mass moments about the COM (quadrupole law)
  float4 *nodeL = nullptr; // {Sxx/M, Sxy/M, Syy/M, log2(M)/16} (learned law)
  bf16x8 *mlpPk = nullptr; // packed forward weights of the learned law
  float *mlpBc = nullptr;  // its biases, accumulator order
  bool mlpLoaded = false;  // nbx_set_mlp has been called
  int4 *nodeI = nullptr;
  float *centerX = nullptr, *centerY = nullptr, *size = nullptr;
  float2 *wsum = nullptr;
  int32_t *bigNodes = nullptr;
  int32_t *subtreeRoots = nullptr;
  uint8_t *nodeDone = nullptr;
  float4 *bodyPack = nullptr; // {x, y, m, 0} in tree order, for the walk
  uint4 *bboxParts = nullptr;
  bool treeBuilt = false;
  bool momentsBuilt = false; // the last build computed nodeQ
  bool learnBuilt = false;   // ... and nodeL
  uint32_t walkBlock = kWalkBlock; // walk workgroup size (nbx_tune_walk)
  bool collided = false;

  Ctl *ctl = nullptr;

  // profiling: NBX_PHASES + 1 events per step, recorded on the stream and
  // read back once per API call (no per-step host sync)
  bool profiling = false;
  std::vector<hipEvent_t> evPool;
  size_t evUsed = 0;
  double phaseMs[NBX_PHASES]{};
  uint64_t profSteps = 0;

  // captured work, by the buffer holding the state when the step starts:
  // the whole step, and each phase on its own (used when profiling, so
  // events can sit between the phases)
  hipGraphExec_t graph[2]{};
  hipGraphExec_t phaseGraph[NBX_PHASES][2]{};
};

namespace {

template <class T>
int devAlloc(nbx_sim *s, T **p, uint64_t count, bool zero = false) {
  void *q = nullptr;
  const size_t bytes = std::max<uint64_t>(count, 1) * sizeof(T);
  NBX_HIP(hipMalloc(&q, bytes));
  s->allocs.push_back(q);
  if (zero) {
    NBX_HIP(hipMemset(q, 0, bytes));
  }
  *p = static_cast<T *>(q);
  return NBX_OK;
}

void dropGraphs(nbx_sim *s) {
  for (auto &g : s->graph) {
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  }
  for (auto &row : s->phaseGraph) {
    for (auto &g : row) {
      if (g) {
        (void)hipGraphExecDestroy(g);
        g = nullptr;
      }
    }
  }
}

int checkLaunch() {
  NBX_HIP(hipGetLastError());
  return NBX_OK;
}

#define LAUNCH(kernel, grid, block, ...)                                      \
  do {                                                                        \
    hipLaunchKernelGGL((kernel), dim3(grid), dim3(block), 0, s->stream,       \
                       __VA_ARGS__);                                          \
    NBX_TRY(checkLaunch());                                                   \
  } while (0)

int launchIterate(nbx_sim *s, bool resetCollide = false);
int launchCollide(nbx_sim *s, bool countersReset = false);
int launchBuild(nbx_sim *s);
int fillKeyLow(nbx_sim *s);
int launchForces(nbx_sim *s, bool count);

// ctl non-null: also resets the collide's counters (the step's collide
// follows).
__global__ __launch_bounds__(kBlock) void k_iterate(
    uint32_t n, float dt, float *__restrict__ px, float *__restrict__ py,
    float *__restrict__ vx, float *__restrict__ vy,
    const float *__restrict__ ax, const float *__restrict__ ay,
    Ctl *__restrict__ ctl) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (ctl && i == 0) {
    resetCollideCounters(ctl);
  }
  if (i >= n) {
    return;
  }
  // Body::update (Body.h:20-23): kick with last step's acceleration, drift.
  const float nvx = vx[i] + ax[i] * dt;
  const float nvy = vy[i] + ay[i] * dt;
  vx[i] = nvx;
  vy[i] = nvy;
  px[i] = px[i] + nvx * dt;
  py[i] = py[i] + nvy * dt;
}

// The step's iterate with the collide's first pass (k_cells) on the moved
// bodies: the positions are binned as they are written. Thread 0 resets the
// collide's counters; the large bodies are counted in largeFused, which no
// other kernel of this launch touches.
__global__ __launch_bounds__(kBlock) void k_iterate_cells(
    uint32_t n, float dt, float *__restrict__ px, float *__restrict__ py,
    float *__restrict__ vx, float *__restrict__ vy,
    const float *__restrict__ ax, const float *__restrict__ ay,
    const float *__restrict__ r, Ctl *__restrict__ ctl, uint32_t mask,
    uint32_t *__restrict__ bucketOf, uint32_t *__restrict__ bucketRank,
    uint32_t *__restrict__ bucketCount, int32_t *__restrict__ largeList,
    float4 *__restrict__ largeBox, int32_t *__restrict__ ufParent) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i == 0) {
    resetCollideCounters(ctl);
  }
  if (i >= n) {
    return;
  }
  // Body::update (Body.h:20-23), as k_iterate.
  const float nvx = vx[i] + ax[i] * dt;
  const float nvy = vy[i] + ay[i] * dt;
  const float nx = px[i] + nvx * dt;
  const float ny = py[i] + nvy * dt;
  vx[i] = nvx;
  vy[i] = nvy;
  px[i] = nx;
  py[i] = ny;
  cellOf(i, nx, ny, r[i], ctl, &ctl->largeFused, mask, bucketOf, bucketRank,
         bucketCount, largeList, largeBox, ufParent);
}

// Buckets of the collide's grid hash for n bodies (a power of two).
uint32_t collideBuckets(uint32_t n) {
  return static_cast<uint32_t>(
      std::max<uint64_t>(1024, pow2AtLeast(2ull * n)));
}

// The step's iterate does the collide's cell pass (k_iterate_cells) when the
// grid collide follows it.
bool cellsInIterate(const nbx_sim *s) {
  return s->params.collisions != 0 &&
         s->params.collide_mode != NBX_COLLIDE_JACOBI && s->n >= 2;
}

__global__ void k_zero_counts(Ctl *__restrict__ ctl) {
  ctl->farCount = 0;
  ctl->bodyPairCount = 0;
  ctl->waveVisits = 0;
  ctl->waveLeafTerms = 0;
  for (int k = 0; k < kWalkStats; ++k) {
    ctl->walkStats[k] = 0;
  }
}

// resetCollide: the step's iterate, whose collide then skips its own
// counter reset.
int launchIterate(nbx_sim *s, bool resetCollide) {
  const uint32_t n = static_cast<uint32_t>(s->n);
  if (n == 0) {
    return NBX_OK;
  }
  const int c = s->cur;
  if (resetCollide && cellsInIterate(s)) {
    LAUNCH(k_iterate_cells, blocksFor(n), kBlock, n, s->params.dt, s->px[c],
           s->py[c], s->vx[c], s->vy[c], s->ax, s->ay, s->r[c], s->ctl,
           collideBuckets(n) - 1, s->bucketOf, s->bucketRank, s->bucketCount,
           s->largeList, s->largeBox, s->ufParent);
    return NBX_OK;
  }
  LAUNCH(k_iterate, blocksFor(n), kBlock, n, s->params.dt, s->px[c], s->py[c],
         s->vx[c], s->vy[c], s->ax, s->ay, resetCollide ? s->ctl : nullptr);
  return NBX_OK;
}

// Exclusive scan of `count` words, clearing the input: one launch of
// k_scan_onelaunch where its grid fits the CUs, else reduce / tile scan /
// down-sweep launches.
int launchScan(nbx_sim *s, uint32_t count, uint32_t *in, uint32_t *out) {
  if (count == 0) {
    return NBX_OK;
  }
  if (s->sortGrid) {
#define SCAN_LAUNCH(ITEMS)                                                    \
  {                                                                           \
    const uint32_t g = blocksFor(count, (ITEMS) * kPsBlock);                  \
    if (g <= s->sortGrid) {                                                   \
      const uint32_t chunk = ((count + g - 1) / g + 3) & ~3u;                 \
      const uint32_t grid = (count + chunk - 1) / chunk;                      \
      LAUNCH((k_scan_onelaunch<ITEMS>), grid, kPsBlock, count, chunk, in, out, \
             s->scanTiles, s->sortBarrier, s->barrierSpins, s->ctl);          \
      return NBX_OK;                                                          \
    }                                                                         \
  }
    SCAN_LAUNCH(8)
    SCAN_LAUNCH(16)
#undef SCAN_LAUNCH
  }
  const uint32_t tiles = blocksFor(count, kScanTile);
  LAUNCH(k_scan_reduce, tiles, kBlock, count, in, s->scanTiles);
  LAUNCH(k_scan_tiles, 1, kBlock, tiles, s->scanTiles);
  LAUNCH(k_scan_down, tiles, kBlock, count, in, s->scanTiles, out);
  return NBX_OK;
}

// One launch of k_sort_onelaunch (one workgroup per 1024 x 1/2/4/8/16 keys,
// the smallest tile that lets at most one workgroup per CU cover n).
// Returns false when no tile fits (or the handle uses three launches a
// pass); build = the tree build's sort with its keys computed in the launch.
bool launchOneLaunchSort(nbx_sim *s, uint32_t n, uint32_t *const keys[2],
                         uint32_t *const vals[2], bool build, int *rc) {
  *rc = NBX_OK;
  if (!s->sortGrid) {
    return false;
  }
  const int c = s->cur;
  const float *px = build ? s->px[c] : nullptr;
  const float *py = build ? s->py[c] : nullptr;
  uint32_t *parts = reinterpret_cast<uint32_t *>(s->bboxParts);
#define SORT_LAUNCH(ITEMS)                                                    \
  {                                                                           \
    const uint32_t g = blocksFor(n, (ITEMS) * kPsBlock);                      \
    if (g <= s->sortGrid) {                                                   \
      const uint32_t chunk = (n + g - 1) / g;                                 \
      if (build) {                                                            \
        hipLaunchKernelGGL((k_sort_onelaunch<ITEMS, true>), dim3(g),          \
                           dim3(kPsBlock), 0, s->stream, n, chunk, keys[0],   \
                           vals[0], keys[1], vals[1], s->sortHist,            \
                           s->sortRowTotal, s->sortBarrier, s->barrierSpins,  \
                           s->ctl, px, py, parts);                            \
      } else {                                                                \
        hipLaunchKernelGGL((k_sort_onelaunch<ITEMS, false>), dim3(g),         \
                           dim3(kPsBlock), 0, s->stream, n, chunk, keys[0],   \
                           vals[0], keys[1], vals[1], s->sortHist,            \
                           s->sortRowTotal, s->sortBarrier, s->barrierSpins,  \
                           s->ctl, px, py, parts);                            \
      }                                                                       \
      *rc = checkLaunch();                                                    \
      return true;                                                            \
    }                                                                         \
  }
  SORT_LAUNCH(1)
  SORT_LAUNCH(2)
  SORT_LAUNCH(4)
  SORT_LAUNCH(8)
  SORT_LAUNCH(16)
#undef SORT_LAUNCH
  return false;
}

// Stable LSD radix sort of 32-bit keys with 32-bit payloads, 4 x 8 bits,
// from buffers 0 to buffers 0 (through 1): one launch where it fits, else
// three launches per pass.
int launchRadixSort(nbx_sim *s, uint32_t n, uint32_t *const keys[2],
                    uint32_t *const vals[2]) {
  if (n == 0) {
    return NBX_OK;
  }
  int rc = NBX_OK;
  if (launchOneLaunchSort(s, n, keys, vals, false, &rc)) {
    return rc;
  }
  const uint32_t tiles = blocksFor(n, kSortTile);
  for (int pass = 0; pass < 4; ++pass) {
    const int src = pass & 1, dst = src ^ 1;
    LAUNCH(k_rs_hist, tiles, kBlock, n, keys[src], 8 * pass, s->sortHist,
           tiles);
    LAUNCH(k_rs_scan, 256, kBlock, tiles, s->sortHist, s->sortRowTotal);
    LAUNCH(k_rs_scatter, tiles, kBlock, n, keys[src], vals[src], 8 * pass,
           s->sortHist, s->sortRowTotal, tiles, keys[dst], vals[dst]);
  }
  return NBX_OK;
}

// The reference GPU collide (Jacobi resolution, k_jacobi.h). Scratch comes
// from buffers the sequential collide and the build use at other times:
// keyHi/vals (the two sorts), keyLo (high key words), grid (sorted 64-bit
// keys, then staged results), bucketOf/bucketRank (per-body contact
// counts, kept for nbx_download_contacts).
int launchCollideJacobi(nbx_sim *s) {
  const uint32_t n = static_cast<uint32_t>(s->n);
  const int c = s->cur;
  float *px = s->px[c], *py = s->py[c], *vx = s->vx[c], *vy = s->vy[c];
  const float *m = s->m[c], *r = s->r[c];
  const uint32_t nb = blocksFor(n);
  // Keys: low words into keyHi[0] (sorted first), high words into keyLo.
  LAUNCH(k_jac_keys, nb, kBlock, n, px, py, r, s->jacobiCell, s->ctl,
         s->keyHi[0], s->keyLo, s->vals[0]);
  NBX_TRY(launchRadixSort(s, n, s->keyHi, s->vals));
  LAUNCH(k_jac_gather_hi, nb, kBlock, n, s->vals[0], s->keyLo, s->keyHi[0]);
  NBX_TRY(launchRadixSort(s, n, s->keyHi, s->vals));
  uint64_t *keysSorted = reinterpret_cast<uint64_t *>(s->grid);
  const uint32_t *indexSorted = s->vals[0];
  float4 *out = s->grid + n; // past the 8 B/body of sorted keys
  int32_t *contacts = reinterpret_cast<int32_t *>(s->bucketOf);
  int32_t *asLater = reinterpret_cast<int32_t *>(s->bucketRank);
  LAUNCH(k_jac_sorted, nb, kBlock, n, px, py, r, indexSorted, s->ctl,
         keysSorted);
  LAUNCH(k_jac_small, nb, kBlock, n, s->ctl, keysSorted, indexSorted, px, py,
         vx, vy, m, r, out, contacts, asLater);
  const uint32_t chunks = blocksFor(n, kJacChunk);
  const uint32_t ys = std::min<uint32_t>(64, std::max<uint32_t>(1, s->jacMaxLarge));
  hipLaunchKernelGGL(k_jac_large_partial, dim3(chunks, ys), dim3(kJacLanes), 0,
                     s->stream, n, chunks, s->jacMaxLarge, s->ctl, indexSorted,
                     px, py, vx, vy, m, r, s->jacPartials);
  NBX_TRY(checkLaunch());
  LAUNCH(k_jac_large_combine, 16, kBlock, n, chunks, s->jacMaxLarge, s->ctl,
         indexSorted, px, py, vx, vy, m, r, s->jacPartials, out, contacts,
         asLater);
  LAUNCH(k_jac_commit, nb, kBlock, n, s->ctl, out, px, py, vx, vy);
  s->jacobiCollided = true;
  s->collided = true;
  return NBX_OK;
}

// countersReset: the step's iterate has reset the collide's counters.
int launchCollide(nbx_sim *s, bool countersReset) {
  const uint32_t n = static_cast<uint32_t>(s->n);
  if (n < 2) {
    return NBX_OK;
  }
  if (s->params.collide_mode == NBX_COLLIDE_JACOBI) {
    return launchCollideJacobi(s);
  }
  s->jacobiCollided = false;
  const int c = s->cur;
  float *px = s->px[c], *py = s->py[c], *vx = s->vx[c], *vy = s->vy[c];
  const float *m = s->m[c], *r = s->r[c];
  const uint32_t nb = blocksFor(n);
  const uint32_t buckets = collideBuckets(n);
  const uint32_t mask = buckets - 1;
  // The step's iterate (k_iterate_cells) has done the cell pass.
  const bool fused = countersReset && cellsInIterate(s);
  const uint32_t pairGrid = std::min<uint32_t>(blocksFor(s->pairCap), 2048);

  if (!countersReset) {
    LAUNCH(k_collide_reset, 1, 1, s->ctl);
  }
  if (!fused) {
    LAUNCH(k_cells, nb, kBlock, n, px, py, r, s->ctl, mask, s->bucketOf,
           s->bucketRank, s->bucketCount, s->largeList, s->largeBox,
           s->ufParent);
  }
  NBX_TRY(launchScan(s, buckets + 1, s->bucketCount, s->bucketStart));
  LAUNCH(k_grid_scatter, nb, kBlock, n, px, py, r, s->ctl, s->bucketOf,
         s->bucketRank, s->bucketStart, s->grid, fused ? 1u : 0u);
  LAUNCH(k_find_small, blocksFor(n, kFindBlock), kFindBlock, n, px, py, r,
         s->ctl, mask, s->bucketOf, s->bucketStart, s->grid, s->largeList,
         s->largeBox, s->pairs, s->pairCap, s->ufParent);
  LAUNCH(k_union, pairGrid, kBlock, s->ctl, static_cast<uint32_t>(n),
         s->pairCap, s->pairs, s->ufParent);
  LAUNCH(k_comp_count, pairGrid, kBlock, s->ctl, s->pairCap, s->pairs,
         s->ufParent, s->pairComp, s->compCount);
  // Component segments: one launch where 4/8/16 bodies per lane cover n
  // with at most one workgroup per CU (block totals in the sort's table,
  // unused during the collide), else one launch with per-workgroup atomics.
  bool allocated = false;
#define ALLOC_LAUNCH(ITEMS)                                                   \
  if (!allocated && s->sortGrid &&                                            \
      blocksFor(n, (ITEMS) * kPsBlock) <= s->sortGrid) {                      \
    const uint32_t g = blocksFor(n, (ITEMS) * kPsBlock);                      \
    const uint32_t chunk = ((n + g - 1) / g + 3) & ~3u;                       \
    LAUNCH((k_comp_alloc_onelaunch<ITEMS>), (n + chunk - 1) / chunk,          \
           kPsBlock, n, chunk, s->ctl, s->compCount, s->compStart,            \
           s->compInfo, s->sortHist, s->sortBarrier, s->barrierSpins);        \
    allocated = true;                                                         \
  }
  ALLOC_LAUNCH(4)
  ALLOC_LAUNCH(8)
  ALLOC_LAUNCH(16)
#undef ALLOC_LAUNCH
  if (!allocated) {
    LAUNCH(k_comp_alloc, blocksFor(n, kBlock * kAllocItems), kBlock, n,
           s->ctl, s->compCount, s->compStart, s->compInfo);
  }
  LAUNCH(k_comp_scatter, pairGrid, kBlock, s->ctl, s->pairCap, s->pairs,
         s->pairComp, s->compCount, s->compStart, s->segPairs);
  LAUNCH(k_resolve_small, nb, kBlock, s->ctl, s->compInfo, s->segPairs,
         s->tierList, s->tierCap, s->bigList, px, py, vx, vy, m, r);
  // Components of 9..1024 pairs, tiers 0..2 in one launch: four
  // workgroups of two waves are resident on each of the 256 CUs (LDS), and
  // one more a CU takes up the slots that tier 2's workgroups free.
  LAUNCH(k_resolve_tiers, 256 * 5, 2 * kWave, s->ctl, s->tierList, s->tierCap,
         s->bigList, s->segPairs, px, py, vx, vy, m, r, 256u * 4);
  // Tier 3 (1025..4096 pairs) and the workgroup path in one launch.
  LAUNCH(k_resolve_large, 256, kBlock, s->ctl, s->tierList, s->tierCap,
         s->compInfo, s->bigList, s->segPairs, s->pairs, s->bodySlot, px, py,
         vx, vy, m, r);
  s->collided = true;
  return NBX_OK;
}

int launchBuild(nbx_sim *s) {
  const uint32_t n = static_cast<uint32_t>(s->n);
  if (n == 0) {
    s->treeBuilt = false;
    return NBX_OK;
  }
  const int c = s->cur, o = c ^ 1;
  const uint32_t nb = blocksFor(n);
  const uint32_t parts = std::min<uint32_t>(nb, kBboxParts);
  // Bounding box, keys and the stable LSD radix sort of the high key words
  // (ending in buffer 0): one launch, or k_bbox, k_keys and the sort's
  // passes.
  int rc = NBX_OK;
  if (launchOneLaunchSort(s, n, s->keyHi, s->vals, true, &rc)) {
    NBX_TRY(rc);
  } else {
    LAUNCH(k_bbox, parts, kBlock, n, s->px[c], s->py[c], s->bboxParts);
    LAUNCH(k_keys, nb, kBlock, n, s->px[c], s->py[c], s->bboxParts, parts,
           s->ctl, s->keyHi[0], s->keyLo, s->vals[0]);
    NBX_TRY(launchRadixSort(s, n, s->keyHi, s->vals));
  }
  const uint32_t *hi = s->keyHi[0];
  uint32_t *perm = s->vals[0];
  // The collide's grid records are free during the build: the permute's
  // staged records (2 float4 a body) when the sort moved many bodies far.
  s->keyLowFilled = false;
  LAUNCH(k_deep, nb, kBlock, n, hi, s->keyLo, perm, s->ctl, s->px[c], s->py[c],
         s->vx[c], s->vy[c], s->m[c], s->r[c], s->id[c], s->grid);
  s->keySource = c;
  LAUNCH(k_permute, nb, kBlock, n, perm, s->px[c], s->py[c], s->vx[c],
         s->vy[c], s->m[c], s->r[c], s->id[c], s->px[o], s->py[o], s->vx[o],
         s->vy[o], s->m[o], s->r[o], s->id[o], s->bodyPack, s->ctl, s->grid);
  s->cur = o;
  const uint32_t tiles256 = nb;
  LAUNCH(k_lcp_count, tiles256, kBlock, n, hi, s->keyLo, perm, tiles256,
         s->tileCount);
  LAUNCH(k_level_scan, kMaxDepth, 1024, tiles256, s->tileCount, s->ctl, n,
         s->nodeCap, s->nodeF, s->nodeI, s->centerX, s->centerY, s->size,
         s->nodeOfSplit);
  LAUNCH(k_emit, tiles256, kBlock, n, hi, s->keyLo, perm, tiles256,
         s->tileCount, s->ctl, s->splitList);
  const uint32_t nodeGrid = std::max<uint32_t>(64, std::min<uint32_t>(nb, 2048));
  LAUNCH(k_nodes, nodeGrid, kBlock, n, hi, s->keyLo, perm, s->ctl,
         s->splitList, s->nodeF, s->nodeI, s->centerX, s->centerY, s->size,
         s->nodeOfSplit);
  // Second mass moments only for the quadrupole and learned laws (null: not
  // computed).
  const bool learned = s->params.law == NBX_LAW_LEARNED;
  float4 *nodeQ =
      (s->params.law == NBX_LAW_ACCURATE_QUAD || learned) ? s->nodeQ : nullptr;
  LAUNCH(k_node_finish, nodeGrid, kBlock, s->ctl, s->px[o], s->py[o], s->m[o],
         s->nodeOfSplit, s->nodeF, s->nodeI, s->wsum, s->nodeDone,
         s->subtreeRoots, nodeQ);
  LAUNCH(k_subtree, std::min<uint32_t>(1024, std::max<uint32_t>(1, n / 1024)),
         kBlock, s->ctl, s->subtreeRoots, s->nodeF, s->nodeI, s->wsum,
         s->nodeDone, nodeQ);
  LAUNCH(k_upward_top, 1, kTopBlock, s->ctl, s->nodeF, s->nodeI, s->wsum,
         s->nodeDone, s->bigNodes, nodeQ);
  if (learned) {
    LAUNCH(k_node_learn, blocksFor(s->nodeCap), kBlock, s->nodeF, s->nodeQ,
           s->ctl, s->nodeL);
  }
  s->treeBuilt = true;
  s->momentsBuilt = nodeQ != nullptr;
  s->learnBuilt = learned;
  return NBX_OK;
}

int launchForces(nbx_sim *s, bool count) {
  const uint32_t n = static_cast<uint32_t>(s->n);
  if (n == 0) {
    return NBX_OK;
  }
  if (!s->treeBuilt) {
    return setError(NBX_ERR_STATE, "forces requested before any tree build");
  }
  if (s->params.law == NBX_LAW_ACCURATE_QUAD && !s->momentsBuilt) {
    return setError(NBX_ERR_STATE,
                    "the quadrupole law needs a tree built under it "
                    "(nbx_build, nbx_attract or nbx_step)");
  }
  if (s->params.law == NBX_LAW_LEARNED && (!s->learnBuilt || !s->mlpLoaded)) {
    return setError(NBX_ERR_STATE,
                    s->mlpLoaded ? "the learned law needs a tree built under it "
                                   "(nbx_build, nbx_attract or nbx_step)"
                                 : "the learned law needs weights (nbx_set_mlp)");
  }
  const int c = s->cur;
  const float tSq = s->params.theta * s->params.theta;
  const float eSq = s->params.epsilon * s->params.epsilon;
  const uint32_t grid = blocksFor(n);
  const bool ref = s->params.summation == NBX_SUM_REFERENCE;
  const bool acc = s->params.law == NBX_LAW_ACCURATE;
  const bool quad = s->params.law == NBX_LAW_ACCURATE_QUAD;
#define WALK(SUMV, LAWV, CNT)                                                 \
  LAUNCH((k_walk<SUMV, LAWV, CNT>), grid, kBlock, n, s->nodeF, s->nodeI,      \
         s->px[c], s->py[c], s->m[c], tSq, eSq, s->ax, s->ay, s->ctl)
  // Reference law: every factor is |m| * fisr(d2) / (d2 + eps2) with
  // |m| <= sum |m| and fisr(d2) <= fisr(0) < 2e19, so with the bound below
  // the min(., FLT_MAX) guard cannot change a value. (The accurate law keeps
  // the guard: its rsqrt(0) is infinite for a body's own term.)
  const double eps2 = static_cast<double>(eSq);
  const bool capFree = !acc && !quad && eps2 > 0 &&
                       s->massAbsSum * 2e19 * std::max(1.0, 1.0 / eps2) < 1e37;
  // The node-mass test is needed only when some body mass is not positive
  // and finite (and always by the counting walk, which counts far nodes of
  // positive mass).
  // theta^2 a power of two: its exponent shift for the scaled opening test
  // of the reference-law walk (-1: not a power of two).
  int tShift = -1;
  {
    int e = 0;
    const float mnt = std::frexp(tSq, &e);
    if (tSq > 0.0f && mnt == 0.5f && e - 1 <= 0 && e - 1 >= -60) {
      tShift = (1 - e) << 23;
    }
  }
#define FAST_B(LAWV, CNT, CAPV, MASSV, BL)                                    \
  LAUNCH((k_walk_fast<LAWV, CNT, CAPV, MASSV, BL>), blocksFor(n, BL), BL, n,  \
         s->nodeF, s->nodeI, s->nodeQ, s->bodyPack, tSq, eSq, s->ax, s->ay,   \
         s->ctl, tShift)
  // Workgroup size: the handle's (nbx_set_walk_block, nbx_tune_walk) for
  // the force walk; the counting walk keeps the default.
#define FAST_MB(LAWV, CNT, CAPV, MASSV)                                       \
  do {                                                                        \
    if constexpr (CNT) {                                                      \
      FAST_B(LAWV, CNT, CAPV, MASSV, kWalkBlock);                             \
    } else if (s->walkBlock == 64) {                                          \
      FAST_B(LAWV, CNT, CAPV, MASSV, 64);                                     \
    } else if (s->walkBlock == 256) {                                         \
      FAST_B(LAWV, CNT, CAPV, MASSV, 256);                                    \
    } else {                                                                  \
      FAST_B(LAWV, CNT, CAPV, MASSV, 128);                                    \
    }                                                                         \
  } while (0)
#define FAST_M(LAWV, CNT, CAPV)                                               \
  do {                                                                        \
    if (!(CNT) && s->massPositive) {                                          \
      FAST_MB(LAWV, CNT, CAPV, false);                                        \
    } else {                                                                  \
      FAST_MB(LAWV, CNT, CAPV, true);                                         \
    }                                                                         \
  } while (0)
#define FAST(LAWV, CNT)                                                       \
  do {                                                                        \
    if (capFree) {                                                            \
      FAST_M(LAWV, CNT, false);                                               \
    } else {                                                                  \
      FAST_M(LAWV, CNT, true);                                                \
    }                                                                         \
  } while (0)
#define LEARNED(CNT, MASSV)                                                   \
  LAUNCH((k_walk_learned<CNT, MASSV>), blocksFor(n, kWalkLearnedBlock),       \
         kWalkLearnedBlock, n, s->nodeF, s->nodeI, s->nodeL, s->centerX,      \
         s->centerY, s->size, s->bodyPack, s->mlpPk, s->mlpBc, tSq, eSq,      \
         s->ax, s->ay, s->ctl)
  if (s->params.law == NBX_LAW_LEARNED) {
    if (count) LEARNED(true, true);
    else if (s->massPositive) LEARNED(false, false);
    else LEARNED(false, true);
  } else if (ref) {
    if (count) {
      if (acc) WALK(kSumReference, kLawAccurate, true);
      else WALK(kSumReference, kLawReference, true);
    } else {
      if (acc) WALK(kSumReference, kLawAccurate, false);
      else WALK(kSumReference, kLawReference, false);
    }
  } else {
    if (count) {
      if (quad) FAST_M(kLawQuadrupole, true, true);
      else if (acc) FAST(kLawAccurate, true);
      else FAST(kLawReference, true);
    } else {
      if (quad) FAST_M(kLawQuadrupole, false, true);
      else if (acc) FAST(kLawAccurate, false);
      else FAST(kLawReference, false);
    }
  }
#undef LEARNED
#undef FAST
#undef FAST_M
#undef FAST_MB
#undef FAST_B
#undef WALK
  return NBX_OK;
}

int record(nbx_sim *s) {
  if (!s->profiling) {
    return NBX_OK;
  }
  if (s->evUsed == s->evPool.size()) {
    hipEvent_t e = nullptr;
    NBX_HIP(hipEventCreate(&e));
    s->evPool.push_back(e);
  }
  NBX_HIP(hipEventRecord(s->evPool[s->evUsed++], s->stream));
  return NBX_OK;
}

// Folds the recorded per-step phase events into the phase totals.
int collectEvents(nbx_sim *s) {
  const size_t per = NBX_PHASES + 1;
  if (s->evUsed < per) {
    s->evUsed = 0;
    return NBX_OK;
  }
  NBX_HIP(hipEventSynchronize(s->evPool[s->evUsed - 1]));
  const size_t steps = s->evUsed / per;
  for (size_t k = 0; k < steps; ++k) {
    for (int p = 0; p < NBX_PHASES; ++p) {
      float ms = 0.0f;
      NBX_HIP(hipEventElapsedTime(&ms, s->evPool[k * per + p],
                                  s->evPool[k * per + p + 1]));
      s->phaseMs[p] += ms;
    }
  }
  s->profSteps += steps;
  s->evUsed = 0;
  return NBX_OK;
}

// A phase of a step: the iterate resets the counters of the collide that
// follows it.
int launchPhase(nbx_sim *s, int phase) {
  switch (phase) {
  case NBX_PHASE_ITERATE:
    return launchIterate(s, s->params.collisions != 0);
  case NBX_PHASE_COLLIDE:
    return s->params.collisions ? launchCollide(s, true) : NBX_OK;
  case NBX_PHASE_BUILD:
    return launchBuild(s);
  default:
    return launchForces(s, false);
  }
}

// One simulation step (Simulation.cpp:11-16), launched eagerly.
int launchStep(nbx_sim *s) {
  for (int p = 0; p < NBX_PHASES; ++p) {
    NBX_TRY(record(s));
    NBX_TRY(launchPhase(s, p));
  }
  return record(s);
}

// Captures `phase` (or the whole step when phase < 0) of a step that starts
// with the state in buffer `start`.
int capture(nbx_sim *s, int phase, int start, hipGraphExec_t *out) {
  const int saved = s->cur;
  const bool savedProf = s->profiling;
  s->cur = (phase == NBX_PHASE_WALK) ? (start ^ 1) : start;
  s->profiling = false;
  hipGraph_t g = nullptr;
  NBX_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  const int rc = phase < 0 ? launchStep(s) : launchPhase(s, phase);
  hipError_t e = hipStreamEndCapture(s->stream, &g);
  s->cur = saved;
  s->profiling = savedProf;
  if (rc != NBX_OK) {
    if (g) {
      (void)hipGraphDestroy(g);
    }
    return rc;
  }
  if (e != hipSuccess) {
    return setError(NBX_ERR_HIP,
                    std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    *out = nullptr;
    return setError(NBX_ERR_HIP,
                    std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  return NBX_OK;
}

// One step from captured graphs: a single graph, or (profiling) one graph
// per phase with events recorded between them.
int graphStep(nbx_sim *s) {
  const int c = s->cur;
  if (!s->profiling) {
    if (!s->graph[c]) {
      NBX_TRY(capture(s, -1, c, &s->graph[c]));
    }
    s->keyLowFilled = false; // the step's build
    NBX_HIP(hipGraphLaunch(s->graph[c], s->stream));
  } else {
    for (int p = 0; p < NBX_PHASES; ++p) {
      NBX_TRY(record(s));
      if (p == NBX_PHASE_COLLIDE && !s->params.collisions) {
        continue;
      }
      if (!s->phaseGraph[p][c]) {
        NBX_TRY(capture(s, p, c, &s->phaseGraph[p][c]));
      }
      s->keyLowFilled = false;
      NBX_HIP(hipGraphLaunch(s->phaseGraph[p][c], s->stream));
    }
    NBX_TRY(record(s));
  }
  s->cur = c ^ 1;
  s->treeBuilt = true;
  s->momentsBuilt = s->params.law == NBX_LAW_ACCURATE_QUAD;
  s->collided = s->params.collisions != 0 && s->n >= 2;
  return NBX_OK;
}

int readStatus(nbx_sim *s, uint32_t *status) {
  NBX_HIP(hipMemcpyAsync(status, &s->ctl->status, sizeof(uint32_t),
                         hipMemcpyDeviceToHost, s->stream));
  NBX_HIP(hipStreamSynchronize(s->stream));
  return NBX_OK;
}

int syncAndCheck(nbx_sim *s) {
  uint32_t status = 0;
  NBX_TRY(readStatus(s, &status));
  NBX_TRY(collectEvents(s));
  if (status & kStatusNodeOverflow) {
    return setError(NBX_ERR_CAPACITY, "tree node capacity exceeded");
  }
  if (status & kStatusPairOverflow) {
    return setError(NBX_ERR_CAPACITY, "collision pair capacity exceeded");
  }
  if (status & kStatusDeepTooLarge) {
    return setError(NBX_ERR_CAPACITY,
                    "more than 4096 bodies share one depth-16 tree cell");
  }
  if (status & kStatusTreeInconsistent) {
    return setError(NBX_ERR_STATE, "internal tree consistency check failed");
  }
  if (status & kStatusJacobiTooManyLarge) {
    return setError(NBX_ERR_CAPACITY,
                    "too many large boxes for the Jacobi collide's partial sums");
  }
  if (status & kStatusBarrierStalled) {
    // Leave the barrier words at rest for the next sort.
    const uint32_t cleared = status & ~kStatusBarrierStalled;
    NBX_HIP(hipMemsetAsync(s->sortBarrier, 0, kBarrierWords * sizeof(uint32_t),
                           s->stream));
    NBX_HIP(hipMemcpyAsync(&s->ctl->status, &cleared, sizeof(uint32_t),
                           hipMemcpyHostToDevice, s->stream));
    NBX_HIP(hipStreamSynchronize(s->stream));
    return setError(NBX_ERR_STATE,
                    "grid barrier timed out (workgroups were not all "
                    "resident)");
  }
  return NBX_OK;
}

int readCtl(nbx_sim *s, Ctl *out) {
  NBX_HIP(hipMemcpyAsync(out, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost,
                         s->stream));
  NBX_HIP(hipStreamSynchronize(s->stream));
  return NBX_OK;
}

template <class T>
int copyOut(nbx_sim *s, T *dst, const T *src, uint64_t count) {
  if (dst && count) {
    NBX_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost,
                           s->stream));
  }
  return NBX_OK;
}

template <class T>
int copyIn(nbx_sim *s, T *dst, const T *src, uint64_t count) {
  if (count) {
    NBX_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice,
                           s->stream));
  }
  return NBX_OK;
}

int zeroDevice(nbx_sim *s, void *dst, uint64_t bytes) {
  if (bytes) {
    NBX_HIP(hipMemsetAsync(dst, 0, bytes, s->stream));
  }
  return NBX_OK;
}

int validParams(const nbx_params *p) {
  if (!(p->dt == p->dt) || !(p->theta >= 0.0f) || !(p->epsilon == p->epsilon)) {
    return setError(NBX_ERR_ARG, "dt/theta/epsilon must be numbers, theta >= 0");
  }
  if (p->law != NBX_LAW_REFERENCE && p->law != NBX_LAW_ACCURATE &&
      p->law != NBX_LAW_ACCURATE_QUAD && p->law != NBX_LAW_LEARNED) {
    return setError(NBX_ERR_ARG, "unknown force law");
  }
  if ((p->law == NBX_LAW_ACCURATE_QUAD || p->law == NBX_LAW_LEARNED) &&
      p->summation != NBX_SUM_COMPENSATED) {
    return setError(NBX_ERR_ARG,
                    "the quadrupole and learned laws need compensated summation");
  }
  if (p->summation != NBX_SUM_REFERENCE &&
      p->summation != NBX_SUM_COMPENSATED) {
    return setError(NBX_ERR_ARG, "unknown summation mode");
  }
  if (p->collide_mode != NBX_COLLIDE_SEQUENTIAL &&
      p->collide_mode != NBX_COLLIDE_JACOBI) {
    return setError(NBX_ERR_ARG, "unknown collide mode");
  }
  return NBX_OK;
}

// Every low key word of the last build (the build stores them only where
// high words tie), from the bodies it keyed: before the download and before
// an upload overwrites those bodies.
int fillKeyLow(nbx_sim *s) {
  if (!s->treeBuilt || s->keyLowFilled || s->n == 0) {
    return NBX_OK;
  }
  const uint32_t n = static_cast<uint32_t>(s->n);
  LAUNCH(k_key_low, blocksFor(n), kBlock, n, s->px[s->keySource],
         s->py[s->keySource], s->ctl, s->keyLo);
  s->keyLowFilled = true;
  return NBX_OK;
}

} // namespace

extern "C" {

const char *nbx_last_error(void) { return g_error.c_str(); }

int nbx_abi_version(void) { return NBX_ABI_VERSION; }

void nbx_default_params(nbx_params *p) {
  if (!p) {
    return;
  }
  std::memset(p, 0, sizeof *p);
  p->dt = 0.05f;
  p->theta = 1.0f;
  p->epsilon = 1.0f;
  p->collisions = 1;
  p->law = NBX_LAW_REFERENCE;
  p->summation = NBX_SUM_COMPENSATED;
  p->use_graph = 1;
}

int nbx_create(uint64_t capacity, nbx_sim **out) {
  const char *mode = std::getenv("NBX_SORT");
  const bool multi = mode && std::strcmp(mode, "multi") == 0;
  return nbx_create_ex(capacity, multi ? NBX_CREATE_NO_GRID_BARRIERS : 0u, out);
}

int nbx_create_ex(uint64_t capacity, uint32_t flags, nbx_sim **out) {
  if (!out) {
    return setError(NBX_ERR_ARG, "out is null");
  }
  *out = nullptr;
  if (capacity == 0 || capacity > (1ull << 27)) {
    return setError(NBX_ERR_ARG, "capacity must be in [1, 2^27]");
  }
  if (flags & ~static_cast<uint32_t>(NBX_CREATE_NO_GRID_BARRIERS)) {
    return setError(NBX_ERR_ARG, "unknown creation flag");
  }
  nbx_sim *s = new nbx_sim();
  s->cap = capacity;
  nbx_default_params(&s->params);
  auto fail = [&](int rc) {
    nbx_destroy(s);
    return rc;
  };
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return setError(NBX_ERR_HIP, "hipStreamCreateWithFlags failed");
  }
  const uint64_t n = capacity;
  int rc = NBX_OK;
  {
    double2 *unused = nullptr;
    if (hipGetDevice(&s->device) != hipSuccess) {
      return fail(setError(NBX_ERR_HIP, "hipGetDevice failed"));
    }
    if ((rc = resolveScratch(s->device, &unused)) != NBX_OK) {
      return fail(rc);
    }
  }
#define A(ptr, count, zero)                                                   \
  if ((rc = devAlloc(s, &(ptr), (count), (zero))) != NBX_OK)                  \
    return fail(rc);
  for (int b = 0; b < 2; ++b) {
    A(s->px[b], n, false) A(s->py[b], n, false) A(s->vx[b], n, false)
    A(s->vy[b], n, false) A(s->m[b], n, false) A(s->r[b], n, false)
    A(s->id[b], n, false)
    A(s->keyHi[b], n, true) A(s->vals[b], n, true)
  }
  A(s->ax, n, true) A(s->ay, n, true)
  s->buckets = std::max<uint64_t>(1024, pow2AtLeast(2 * n));
  A(s->bucketOf, n, false) A(s->bucketRank, n, false)
  A(s->bucketCount, s->buckets + 1, true)
  A(s->bucketStart, s->buckets + 1, false)
  A(s->scanTiles, blocksFor(s->buckets + 1, kScanTile) + 1, false)
  A(s->grid, 2 * n, false)
  A(s->largeList, n, false) A(s->ufParent, n, false) A(s->largeBox, n, false)
  s->pairCap = static_cast<uint32_t>(std::min<uint64_t>(8 * n + 1024, 0xffffff00ull));
  A(s->pairs, s->pairCap, false) A(s->segPairs, s->pairCap, false)
  A(s->pairComp, s->pairCap, false)
  A(s->compCount, n, true) A(s->compStart, n, false) A(s->compInfo, n, false)
  A(s->bigList, n, false)
  // A queued component has more than kSmallComp pairs.
  s->tierCap = s->pairCap / (kSmallComp + 1) + 1;
  A(s->tierList, static_cast<uint64_t>(kTierQueues * kWaveTiers) * s->tierCap, false)
  if ((rc = devAlloc(s, &s->bodySlot, n, false)) != NBX_OK ||
      hipMemset(s->bodySlot, 0xff, n * sizeof(int32_t)) != hipSuccess) {
    return fail(rc != NBX_OK ? rc : setError(NBX_ERR_HIP, "hipMemset failed"));
  }
  A(s->keyLo, n, false)
  {
    // The one-launch sort needs every one of its workgroups resident at
    // once: at most one per CU, and only where one fits a CU.
    int dev = 0, cus = 0, perCu = 0;
    // Test hook: a small poll bound forces the barrier-timeout path.
    if (const char *sp = std::getenv("NBX_BARRIER_SPINS")) {
      s->barrierSpins = static_cast<uint32_t>(std::strtoul(sp, nullptr, 10));
    }
    const bool multi = (flags & NBX_CREATE_NO_GRID_BARRIERS) != 0;
    if (!multi && hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                              dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &perCu, reinterpret_cast<const void *>(k_sort_onelaunch<16, true>),
            kPsBlock, 0) == hipSuccess &&
        perCu >= 1 && cus >= 1) {
      s->sortGrid = std::min<uint32_t>(static_cast<uint32_t>(cus), kPsMaxGrid);
    }
  }
  A(s->sortHist,
    std::max<uint64_t>(256ull * blocksFor(n, kSortTile), 256ull * s->sortGrid),
    false)
  A(s->sortRowTotal, 256, false)
  A(s->sortBarrier, kBarrierWords, true)
  A(s->tileCount, static_cast<uint64_t>(kMaxDepth) * blocksFor(n), false)
  s->nodeCap = static_cast<int64_t>(8 * n + 8);
  A(s->splitList, s->nodeCap / 4 + 1, false)
  A(s->nodeOfSplit, s->nodeCap / 4 + 1, false)
  A(s->nodeF, s->nodeCap, false) A(s->nodeI, s->nodeCap, false)
  A(s->nodeQ, s->nodeCap, true)
  A(s->nodeL, s->nodeCap, true)
  A(s->mlpPk, kPkFwd, true)
  A(s->mlpBc, kBcAll, true)
  A(s->centerX, s->nodeCap, false) A(s->centerY, s->nodeCap, false)
  A(s->size, s->nodeCap, false) A(s->wsum, s->nodeCap, false)
  A(s->bigNodes, s->nodeCap / 4 + 1, false)
  A(s->subtreeRoots, s->nodeCap, false)
  A(s->nodeDone, s->nodeCap, false)
  A(s->bodyPack, n + kBodyPad, true)
  A(s->bboxParts, std::max<uint32_t>(kBboxParts, s->sortGrid), false)
  A(s->ctl, 1, true)
  {
    const uint64_t chunks = blocksFor(n, kJacChunk);
    const uint64_t entries = kJacPartialBytes / sizeof(JacAcc);
    s->jacMaxLarge = static_cast<uint32_t>(std::max<uint64_t>(1, entries / chunks));
    A(s->jacPartials, s->jacMaxLarge * chunks, false)
  }
#undef A
  if (hipDeviceSynchronize() != hipSuccess) {
    return fail(setError(NBX_ERR_HIP, "hipDeviceSynchronize failed"));
  }
  *out = s;
  return NBX_OK;
}

int nbx_destroy(nbx_sim *s) {
  if (!s) {
    return NBX_OK;
  }
  if (s->stream) {
    (void)hipStreamSynchronize(s->stream);
  }
  dropGraphs(s);
  for (auto &e : s->evPool) {
    (void)hipEventDestroy(e);
  }
  for (void *p : s->allocs) {
    (void)hipFree(p);
  }
  if (s->stream) {
    (void)hipStreamDestroy(s->stream);
  }
  delete s;
  return NBX_OK;
}

int nbx_upload(nbx_sim *s, uint64_t n, const float *px, const float *py,
               const float *vx, const float *vy, const float *ax,
               const float *ay, const float *m, const float *r,
               const uint32_t *id) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  if (n > s->cap) {
    return setError(NBX_ERR_ARG, "n exceeds the handle's capacity");
  }
  if (n && (!px || !py || !vx || !vy || !m || !r || !id)) {
    return setError(NBX_ERR_ARG, "null body array");
  }
  NBX_HIP(hipStreamSynchronize(s->stream));
  dropGraphs(s);
  NBX_TRY(fillKeyLow(s));
  NBX_HIP(hipStreamSynchronize(s->stream));
  s->cur = 0;
  s->n = n;
  s->massAbsSum = 0;
  s->massPositive = true;
  for (uint64_t k = 0; k < n; ++k) {
    s->massAbsSum += std::fabs(static_cast<double>(m[k]));
    s->massPositive = s->massPositive && m[k] > 0.0f && m[k] <= FLT_MAX;
  }
  // CudaCollider's cell (CudaCollision.cu:197-209): 2 x median radius x 1.01.
  s->jacobiCell = 1.0f;
  // The sequential collide's grid cell: 1.01 x the 5/8 quantile of the box
  // widths 2r (a performance choice only: the pair set does not depend on
  // it), and 1 where that is not positive (Collision.cpp:178-180).
  float gridCell = 1.0f;
  if (n > 0) {
    std::vector<float> radii(r, r + n);
    std::nth_element(radii.begin(), radii.begin() + n / 2, radii.end());
    s->jacobiCell = 2.0f * radii[n / 2] * 1.01f;
    const size_t q = std::min<size_t>(n - 1, (5 * n) / 8);
    std::nth_element(radii.begin(), radii.begin() + q, radii.end());
    gridCell = 2.0f * radii[q] * 1.01f;
  }
  if (!(s->jacobiCell > 0.0f)) {
    s->jacobiCell = 1.0f;
  }
  if (!(gridCell > 0.0f) || !(gridCell <= FLT_MAX)) {
    gridCell = 1.0f;
  }
  s->gridCell[0] = gridCell;
  s->gridCell[1] = 1.0f / gridCell;
  NBX_HIP(hipMemcpyAsync(reinterpret_cast<char *>(s->ctl) + offsetof(Ctl, cell),
                         s->gridCell, sizeof s->gridCell, hipMemcpyHostToDevice,
                         s->stream));
  s->jacobiCollided = false;
  NBX_TRY(copyIn(s, s->px[0], px, n));
  NBX_TRY(copyIn(s, s->py[0], py, n));
  NBX_TRY(copyIn(s, s->vx[0], vx, n));
  NBX_TRY(copyIn(s, s->vy[0], vy, n));
  NBX_TRY(copyIn(s, s->m[0], m, n));
  NBX_TRY(copyIn(s, s->r[0], r, n));
  NBX_TRY(copyIn(s, s->id[0], id, n));
  if (ax) {
    NBX_TRY(copyIn(s, s->ax, ax, n));
  } else {
    NBX_TRY(zeroDevice(s, s->ax, n * sizeof(float)));
  }
  if (ay) {
    NBX_TRY(copyIn(s, s->ay, ay, n));
  } else {
    NBX_TRY(zeroDevice(s, s->ay, n * sizeof(float)));
  }
  NBX_HIP(hipStreamSynchronize(s->stream));
  s->frame = 0;
  s->treeBuilt = false;
  s->collided = false;
  return NBX_OK;
}

int nbx_init_disc(nbx_sim *s, uint64_t n, uint32_t seed) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  if (n > s->cap) {
    return setError(NBX_ERR_ARG, "n exceeds the handle's capacity");
  }
  std::vector<float> px(n), py(n), vx(n), vy(n), m(n), r(n);
  std::vector<uint32_t> id(n);
  nbx_host::uniformDisc(n, seed, px.data(), py.data(), vx.data(), vy.data(),
                        m.data(), r.data(), id.data());
  return nbx_upload(s, n, px.data(), py.data(), vx.data(), vy.data(), nullptr,
                    nullptr, m.data(), r.data(), id.data());
}

int nbx_init_scenario(nbx_sim *s, int scenario, uint64_t n, uint32_t seed) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  if (n > s->cap) {
    return setError(NBX_ERR_ARG, "n exceeds the handle's capacity");
  }
  using Gen = void (*)(uint64_t, uint32_t, float *, float *, float *, float *,
                       float *, float *, uint32_t *);
  Gen gen = nullptr;
  switch (scenario) {
  case NBX_SCENARIO_DISC:
    gen = nbx_host::uniformDisc;
    break;
  case NBX_SCENARIO_TWO_GALAXIES:
    gen = nbx_host::twoGalaxies;
    break;
  case NBX_SCENARIO_COLD_SQUARE:
    gen = nbx_host::coldSquare;
    break;
  case NBX_SCENARIO_MERGER:
    gen = nbx_host::galaxyMerger;
    break;
  default:
    return setError(NBX_ERR_ARG, "unknown scenario");
  }
  if (n < 2 && (scenario == NBX_SCENARIO_TWO_GALAXIES ||
                scenario == NBX_SCENARIO_MERGER)) {
    return setError(NBX_ERR_ARG, "a two-disc scenario needs n >= 2");
  }
  std::vector<float> px(n), py(n), vx(n), vy(n), m(n), r(n);
  std::vector<uint32_t> id(n);
  gen(n, seed, px.data(), py.data(), vx.data(), vy.data(), m.data(), r.data(),
      id.data());
  return nbx_upload(s, n, px.data(), py.data(), vx.data(), vy.data(), nullptr,
                    nullptr, m.data(), r.data(), id.data());
}

int nbx_download(nbx_sim *s, float *px, float *py, float *vx, float *vy,
                 float *ax, float *ay, float *m, float *r, uint32_t *id) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  const int c = s->cur;
  const uint64_t n = s->n;
  NBX_TRY(copyOut(s, px, s->px[c], n));
  NBX_TRY(copyOut(s, py, s->py[c], n));
  NBX_TRY(copyOut(s, vx, s->vx[c], n));
  NBX_TRY(copyOut(s, vy, s->vy[c], n));
  NBX_TRY(copyOut(s, ax, s->ax, n));
  NBX_TRY(copyOut(s, ay, s->ay, n));
  NBX_TRY(copyOut(s, m, s->m[c], n));
  NBX_TRY(copyOut(s, r, s->r[c], n));
  NBX_TRY(copyOut(s, id, s->id[c], n));
  return syncAndCheck(s);
}

int nbx_download_by_id(nbx_sim *s, float *px, float *py, float *vx, float *vy,
                       float *m, float *r) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  const uint64_t n = s->n;
  std::vector<uint32_t> id(n);
  std::vector<float> tmp(n);
  NBX_TRY(nbx_download(s, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr, nullptr, nullptr, id.data()));
  for (uint64_t i = 0; i < n; ++i) {
    if (id[i] >= n) {
      return setError(NBX_ERR_STATE, "ids are not a permutation of 0..n-1");
    }
  }
  float *outs[6] = {px, py, vx, vy, m, r};
  const int c = s->cur;
  float *srcs[6] = {s->px[c], s->py[c], s->vx[c], s->vy[c], s->m[c], s->r[c]};
  for (int f = 0; f < 6; ++f) {
    if (!outs[f]) {
      continue;
    }
    NBX_TRY(copyOut(s, tmp.data(), srcs[f], n));
    NBX_HIP(hipStreamSynchronize(s->stream));
    for (uint64_t i = 0; i < n; ++i) {
      outs[f][id[i]] = tmp[i];
    }
  }
  return NBX_OK;
}

int nbx_count(nbx_sim *s, uint64_t *n) {
  if (!s || !n) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  *n = s->n;
  return NBX_OK;
}

int nbx_set_params(nbx_sim *s, const nbx_params *p) {
  if (!s || !p) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  NBX_TRY(validParams(p));
  if (std::memcmp(&s->params, p, sizeof *p) != 0) {
    NBX_HIP(hipStreamSynchronize(s->stream));
    dropGraphs(s);
    s->params = *p;
  }
  return NBX_OK;
}

int nbx_get_params(nbx_sim *s, nbx_params *p) {
  if (!s || !p) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  *p = s->params;
  return NBX_OK;
}

int nbx_set_walk_block(nbx_sim *s, uint32_t lanes) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  if (lanes != 64 && lanes != 128 && lanes != 256) {
    return setError(NBX_ERR_ARG, "walk workgroup size must be 64, 128 or 256");
  }
  if (lanes != s->walkBlock) {
    NBX_HIP(hipStreamSynchronize(s->stream));
    dropGraphs(s);
    s->walkBlock = lanes;
  }
  return NBX_OK;
}

int nbx_get_walk_block(nbx_sim *s, uint32_t *lanes) {
  if (!s || !lanes) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  *lanes = s->walkBlock;
  return NBX_OK;
}

int nbx_tune_walk(nbx_sim *s, uint32_t reps, uint32_t *best, float *ms) {
  if (!s || !best) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  if (reps == 0 || reps > 1000) {
    return setError(NBX_ERR_ARG, "reps must be in 1..1000");
  }
  if (s->params.summation == NBX_SUM_REFERENCE ||
      s->params.law == NBX_LAW_LEARNED) {
    return setError(NBX_ERR_STATE, "the walk workgroup size applies to the "
                                   "compensated walk of the closed-form laws");
  }
  static constexpr uint32_t kCand[3] = {64, 128, 256};
  if (s->n == 0) {
    *best = s->walkBlock;
    if (ms) {
      std::fill(ms, ms + 3, 0.0f);
    }
    return NBX_OK;
  }
  if (!s->treeBuilt) {
    NBX_TRY(launchBuild(s));
  }
  hipEvent_t ev[2]{};
  NBX_HIP(hipEventCreate(&ev[0]));
  if (hipEventCreate(&ev[1]) != hipSuccess) {
    (void)hipEventDestroy(ev[0]);
    return setError(NBX_ERR_HIP, "hipEventCreate failed");
  }
  const uint32_t prev = s->walkBlock;
  float t[3] = {0, 0, 0};
  int rc = NBX_OK;
  for (int i = 0; i < 3 && rc == NBX_OK; ++i) {
    s->walkBlock = kCand[i];
    rc = launchForces(s, false); // warm-up: code load, caches
    for (uint32_t r = 0; r < reps && rc == NBX_OK; ++r) {
      if (r == 0 && hipEventRecord(ev[0], s->stream) != hipSuccess) {
        rc = setError(NBX_ERR_HIP, "hipEventRecord failed");
        break;
      }
      rc = launchForces(s, false);
    }
    if (rc == NBX_OK) {
      if (hipEventRecord(ev[1], s->stream) != hipSuccess ||
          hipEventSynchronize(ev[1]) != hipSuccess ||
          hipEventElapsedTime(&t[i], ev[0], ev[1]) != hipSuccess) {
        rc = setError(NBX_ERR_HIP, "timing the walk failed");
      }
      t[i] /= static_cast<float>(reps);
    }
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (rc != NBX_OK) {
    s->walkBlock = prev;
    return rc;
  }
  int pick = 0;
  for (int i = 1; i < 3; ++i) {
    if (t[i] < t[pick]) {
      pick = i;
    }
  }
  s->walkBlock = kCand[pick];
  if (s->walkBlock != prev) {
    dropGraphs(s);
  }
  *best = s->walkBlock;
  if (ms) {
    std::copy(t, t + 3, ms);
  }
  return syncAndCheck(s);
}

int nbx_step(nbx_sim *s, uint32_t steps) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  const bool graphs = s->params.use_graph && s->n > 0;
  for (uint32_t k = 0; k < steps; ++k) {
    if (graphs) {
      NBX_TRY(graphStep(s));
    } else {
      NBX_TRY(launchStep(s));
    }
    s->frame += 1;
  }
  return syncAndCheck(s);
}

int nbx_step_async(nbx_sim *s, uint32_t steps) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  if (s->sortGrid) {
    // Its one-launch kernels wait on grid barriers, which need every
    // workgroup resident: not promised while other streams hold CUs.
    return setError(NBX_ERR_STATE,
                    "nbx_step_async needs a handle created with "
                    "NBX_CREATE_NO_GRID_BARRIERS");
  }
  const bool graphs = s->params.use_graph && s->n > 0;
  for (uint32_t k = 0; k < steps; ++k) {
    if (graphs) {
      NBX_TRY(graphStep(s));
    } else {
      NBX_TRY(launchStep(s));
    }
    s->frame += 1;
  }
  return NBX_OK;
}

int nbx_iterate(nbx_sim *s) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  NBX_TRY(launchIterate(s));
  return syncAndCheck(s);
}

int nbx_collide(nbx_sim *s) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  NBX_TRY(launchCollide(s));
  return syncAndCheck(s);
}

int nbx_build(nbx_sim *s) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  NBX_TRY(launchBuild(s));
  return syncAndCheck(s);
}

int nbx_forces(nbx_sim *s) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  NBX_TRY(launchForces(s, false));
  return syncAndCheck(s);
}

int nbx_attract(nbx_sim *s) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  NBX_TRY(launchBuild(s));
  NBX_TRY(launchForces(s, false));
  return syncAndCheck(s);
}

int nbx_frame(nbx_sim *s, uint64_t *frame) {
  if (!s || !frame) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  *frame = s->frame;
  return NBX_OK;
}

int nbx_sync(nbx_sim *s) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  return syncAndCheck(s);
}

int nbx_tree_info(nbx_sim *s, int64_t *nodes, int32_t *levels) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  NBX_TRY(syncAndCheck(s));
  Ctl c{};
  NBX_TRY(readCtl(s, &c));
  if (nodes) {
    *nodes = s->treeBuilt ? c.nodeCount : 0;
  }
  if (levels) {
    *levels = s->treeBuilt ? c.levels : 0;
  }
  return NBX_OK;
}

int nbx_download_tree(nbx_sim *s, int64_t cap, float *comX, float *comY,
                      float *mass, float *centerX, float *centerY,
                      float *size, int32_t *children, int32_t *next,
                      int32_t *start, int32_t *end) {
  int64_t nodes = 0;
  NBX_TRY(nbx_tree_info(s, &nodes, nullptr));
  const uint64_t k = static_cast<uint64_t>(std::min(nodes, cap));
  std::vector<float4> f(k);
  std::vector<int4> iv(k);
  NBX_TRY(copyOut(s, f.data(), s->nodeF, k));
  NBX_TRY(copyOut(s, iv.data(), s->nodeI, k));
  NBX_TRY(copyOut(s, centerX, s->centerX, k));
  NBX_TRY(copyOut(s, centerY, s->centerY, k));
  NBX_TRY(copyOut(s, size, s->size, k));
  NBX_HIP(hipStreamSynchronize(s->stream));
  for (uint64_t i = 0; i < k; ++i) {
    if (comX) comX[i] = f[i].x;
    if (comY) comY[i] = f[i].y;
    if (mass) mass[i] = f[i].z;
    if (children) children[i] = iv[i].x;
    if (next) next[i] = iv[i].y;
    if (start) start[i] = iv[i].z;
    if (end) end[i] = iv[i].w;
  }
  return NBX_OK;
}

int nbx_root_quad(nbx_sim *s, float quad[3]) {
  if (!s || !quad) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  Ctl c{};
  NBX_TRY(readCtl(s, &c));
  quad[0] = c.quad[0];
  quad[1] = c.quad[1];
  quad[2] = c.quad[2];
  return NBX_OK;
}

int nbx_download_keys(nbx_sim *s, uint32_t *hi, uint32_t *lo, uint32_t *perm) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  const uint64_t n = s->n;
  std::vector<uint32_t> p(n), l(n);
  if (lo) {
    NBX_TRY(fillKeyLow(s));
  }
  NBX_TRY(copyOut(s, hi, s->keyHi[0], n));
  NBX_TRY(copyOut(s, p.data(), s->vals[0], n));
  NBX_TRY(copyOut(s, l.data(), s->keyLo, n));
  NBX_HIP(hipStreamSynchronize(s->stream));
  for (uint64_t k = 0; k < n; ++k) {
    if (perm) perm[k] = p[k];
    if (lo) lo[k] = p[k] < n ? l[p[k]] : 0u;
  }
  return NBX_OK;
}

int nbx_collide_info(nbx_sim *s, float *cell, int64_t *pairs,
                     int64_t *components, int64_t *largest) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  NBX_TRY(syncAndCheck(s));
  Ctl c{};
  NBX_TRY(readCtl(s, &c));
  const bool ok = s->collided && s->n >= 2;
  if (ok && s->jacobiCollided) {
    // CudaCollider::lastAabbPairs and its per-body contact counts.
    std::vector<int32_t> cnt(s->n), later(s->n);
    NBX_TRY(copyOut(s, cnt.data(), reinterpret_cast<int32_t *>(s->bucketOf), s->n));
    NBX_TRY(copyOut(s, later.data(), reinterpret_cast<int32_t *>(s->bucketRank), s->n));
    int64_t p = 0, touched = 0, most = 0;
    for (uint64_t i = 0; i < s->n; ++i) {
      p += later[i];
      touched += cnt[i] > 0;
      most = std::max<int64_t>(most, cnt[i]);
    }
    if (cell) *cell = c.cell;
    if (pairs) *pairs = p;
    if (components) *components = touched;
    if (largest) *largest = most;
    return NBX_OK;
  }
  if (cell) *cell = ok ? c.cell : 0.0f;
  if (pairs) *pairs = ok ? c.pairCount : 0;
  if (components) *components = ok ? c.compCount : 0;
  if (largest) *largest = ok ? c.largest : 0;
  return NBX_OK;
}

int nbx_download_contacts(nbx_sim *s, int32_t *contacts, int32_t *aabb_as_later) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  NBX_TRY(syncAndCheck(s));
  if (!s->jacobiCollided) {
    return setError(NBX_ERR_STATE, "the last collide did not run in the Jacobi mode");
  }
  NBX_TRY(copyOut(s, contacts, reinterpret_cast<int32_t *>(s->bucketOf), s->n));
  NBX_TRY(copyOut(s, aabb_as_later, reinterpret_cast<int32_t *>(s->bucketRank), s->n));
  return NBX_OK;
}

int nbx_download_pairs(nbx_sim *s, int64_t cap, int32_t *later,
                       int32_t *earlier, int32_t *comp, int64_t *count) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  NBX_TRY(syncAndCheck(s));
  Ctl c{};
  NBX_TRY(readCtl(s, &c));
  if (!s->collided || s->n < 2) {
    if (count) *count = 0;
    return NBX_OK;
  }
  if (s->jacobiCollided) {
    return setError(NBX_ERR_STATE,
                    "the Jacobi collide does not list pairs (see nbx_download_contacts)");
  }
  const uint64_t total = c.compPairs;
  std::vector<int4> info(c.compCount);
  std::vector<int4> seg(total);
  NBX_TRY(copyOut(s, info.data(), s->compInfo, info.size()));
  NBX_TRY(copyOut(s, seg.data(), s->segPairs, seg.size()));
  NBX_HIP(hipStreamSynchronize(s->stream));
  int64_t k = 0;
  for (const int4 &ci : info) {
    for (int q = 0; q < ci.z; ++q, ++k) {
      if (k < cap) {
        const int4 &p = seg[ci.y + q];
        if (later) later[k] = p.x;
        if (earlier) earlier[k] = p.y;
        if (comp) comp[k] = ci.x;
      }
    }
  }
  if (count) *count = k;
  return NBX_OK;
}

int nbx_count_interactions(nbx_sim *s, uint64_t *far_nodes,
                           uint64_t *body_pairs, uint64_t *wave_visits,
                           uint64_t *wave_leaf_terms) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(1), 0, s->stream, s->ctl);
  NBX_TRY(checkLaunch());
  // Counting walk writes the same accelerations as the plain walk.
  NBX_TRY(launchForces(s, true));
  NBX_TRY(syncAndCheck(s));
  Ctl c{};
  NBX_TRY(readCtl(s, &c));
  if (far_nodes) *far_nodes = c.farCount;
  if (body_pairs) *body_pairs = c.bodyPairCount;
  if (wave_visits) *wave_visits = c.waveVisits;
  if (wave_leaf_terms) *wave_leaf_terms = c.waveLeafTerms;
  return NBX_OK;
}

int nbx_walk_profile(nbx_sim *s, uint64_t *out, uint32_t count) {
  if (!s || (count && !out)) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  if (s->params.summation == NBX_SUM_REFERENCE || s->params.law == NBX_LAW_LEARNED) {
    return setError(NBX_ERR_STATE, "the work profile covers the compensated walk "
                                   "of the closed-form laws");
  }
  hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(1), 0, s->stream, s->ctl);
  NBX_TRY(checkLaunch());
  NBX_TRY(launchForces(s, true));
  NBX_TRY(syncAndCheck(s));
  Ctl c{};
  NBX_TRY(readCtl(s, &c));
  for (uint32_t k = 0; k < count && k < static_cast<uint32_t>(kWalkStats); ++k) {
    out[k] = c.walkStats[k];
  }
  return NBX_OK;
}

int nbx_set_profiling(nbx_sim *s, int on) {
  if (!s) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  NBX_TRY(syncAndCheck(s));
  s->profiling = on != 0;
  return NBX_OK;
}

int nbx_phase_times(nbx_sim *s, double ms[NBX_PHASES], uint64_t *steps) {
  if (!s || !ms) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  NBX_TRY(syncAndCheck(s));
  for (int p = 0; p < NBX_PHASES; ++p) {
    ms[p] = s->phaseMs[p];
    s->phaseMs[p] = 0.0;
  }
  if (steps) *steps = s->profSteps;
  s->profSteps = 0;
  return NBX_OK;
}

int nbx_stream(nbx_sim *s, void **stream) {
  if (!s || !stream) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  *stream = s->stream;
  return NBX_OK;
}

int nbx_energy(nbx_sim *s, double eps2, double *kinetic, double *potential,
               double momentum[2]) {
  if (!s || !kinetic || !potential || !momentum) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  if (!(eps2 >= 0.0)) {
    return setError(NBX_ERR_ARG, "eps2 must be >= 0");
  }
  *kinetic = *potential = momentum[0] = momentum[1] = 0.0;
  const uint64_t n = s->n;
  if (n == 0) {
    return NBX_OK;
  }
  NBX_TRY(syncAndCheck(s));
  const int c = s->cur;
  std::vector<float> m(n), vx(n), vy(n);
  NBX_TRY(copyOut(s, m.data(), s->m[c], n));
  NBX_TRY(copyOut(s, vx.data(), s->vx[c], n));
  NBX_TRY(copyOut(s, vy.data(), s->vy[c], n));
  double *part = nullptr, *phi = nullptr;
  if (hipMalloc(&part, static_cast<size_t>(kExactChunks) * n * sizeof(double)) !=
          hipSuccess ||
      hipMalloc(&phi, n * sizeof(double)) != hipSuccess) {
    (void)hipFree(part);
    (void)hipFree(phi);
    return setError(NBX_ERR_HIP, "hipMalloc failed in nbx_energy");
  }
  const uint32_t nn = static_cast<uint32_t>(n);
  hipLaunchKernelGGL(k_energy_partial, dim3(blocksFor(nn), kExactChunks),
                     dim3(kBlock), 0, s->stream, nn, s->px[c], s->py[c],
                     s->m[c], std::sqrt(eps2), part);
  hipLaunchKernelGGL(k_energy_reduce, dim3(blocksFor(nn)), dim3(kBlock), 0,
                     s->stream, nn, part, phi);
  std::vector<double> hphi(n);
  int rc = NBX_OK;
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(hphi.data(), phi, n * sizeof(double), hipMemcpyDeviceToHost,
                     s->stream) != hipSuccess ||
      hipStreamSynchronize(s->stream) != hipSuccess) {
    rc = setError(NBX_ERR_HIP, "energy kernels failed");
  }
  (void)hipFree(part);
  (void)hipFree(phi);
  NBX_TRY(rc);
  double k = 0.0, u = 0.0, p0 = 0.0, p1 = 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    const double mi = m[i], ux = vx[i], uy = vy[i];
    k += 0.5 * mi * (ux * ux + uy * uy);
    u += 0.5 * mi * hphi[i];
    p0 += mi * ux;
    p1 += mi * uy;
  }
  *kinetic = k;
  *potential = u;
  momentum[0] = p0;
  momentum[1] = p1;
  return NBX_OK;
}

int nbx_exact_forces(nbx_sim *s, uint64_t n_targets, const uint32_t *targets,
                     double eps2, double *ax, double *ay) {
  if (!s || (n_targets && (!targets || !ax || !ay))) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  if (n_targets == 0 || s->n == 0) {
    return NBX_OK;
  }
  for (uint64_t t = 0; t < n_targets; ++t) {
    if (targets[t] >= s->n) {
      return setError(NBX_ERR_ARG, "target index out of range");
    }
  }
  NBX_TRY(syncAndCheck(s));
  uint32_t *dT = nullptr;
  double2 *part = nullptr;
  double *dax = nullptr, *day = nullptr;
  const uint32_t nt = static_cast<uint32_t>(n_targets);
  int rc = NBX_OK;
  auto cleanup = [&]() {
    (void)hipFree(dT);
    (void)hipFree(part);
    (void)hipFree(dax);
    (void)hipFree(day);
  };
  if (hipMalloc(&dT, nt * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&part, static_cast<size_t>(kExactChunks) * nt * sizeof(double2)) !=
          hipSuccess ||
      hipMalloc(&dax, nt * sizeof(double)) != hipSuccess ||
      hipMalloc(&day, nt * sizeof(double)) != hipSuccess) {
    cleanup();
    return setError(NBX_ERR_HIP, "hipMalloc failed in nbx_exact_forces");
  }
  const int c = s->cur;
  if (hipMemcpyAsync(dT, targets, nt * sizeof(uint32_t), hipMemcpyHostToDevice,
                     s->stream) != hipSuccess) {
    rc = setError(NBX_ERR_HIP, "target upload failed");
  }
  if (rc == NBX_OK) {
    hipLaunchKernelGGL(k_exact_partial, dim3(blocksFor(nt), kExactChunks),
                       dim3(kBlock), 0, s->stream,
                       static_cast<uint32_t>(s->n), nt, dT, s->px[c], s->py[c],
                       s->m[c], eps2, part);
    hipLaunchKernelGGL(k_exact_reduce, dim3(blocksFor(nt)), dim3(kBlock), 0,
                       s->stream, nt, part, dax, day);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(ax, dax, nt * sizeof(double), hipMemcpyDeviceToHost,
                       s->stream) != hipSuccess ||
        hipMemcpyAsync(ay, day, nt * sizeof(double), hipMemcpyDeviceToHost,
                       s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess) {
      rc = setError(NBX_ERR_HIP, "exact-force kernels failed");
    }
  }
  cleanup();
  return rc;
}

int nbx_exact_forces_noalloc(nbx_sim *s, uint64_t n_targets,
                             const uint32_t *targets, double eps2, double *ax,
                             double *ay) {
  if (!s || (n_targets && (!targets || !ax || !ay))) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  if (n_targets > kNoAllocTargets) {
    return setError(NBX_ERR_ARG, "nbx_exact_forces_noalloc takes at most " +
                                     std::to_string(kNoAllocTargets) +
                                     " targets");
  }
  if (n_targets == 0 || s->n == 0) {
    return NBX_OK;
  }
  for (uint64_t t = 0; t < n_targets; ++t) {
    if (targets[t] >= s->n) {
      return setError(NBX_ERR_ARG, "target index out of range");
    }
  }
  double2 *base = nullptr;
  NBX_TRY(resolveScratch(s->device, &base));
  NBX_TRY(syncAndCheck(s));
  // Every (chunk, target) partial the reduction reads is written by the
  // partial-sum kernel of this call, so nothing needs clearing first.
  double2 *part = base;
  double *dax = reinterpret_cast<double *>(base + kScratchPartials);
  double *day = dax + kNoAllocTargets;
  uint32_t *dT = reinterpret_cast<uint32_t *>(base + kScratchPartials + kNoAllocTargets);
  const uint32_t nt = static_cast<uint32_t>(n_targets);
  const int c = s->cur;
  if (hipMemcpyAsync(dT, targets, nt * sizeof(uint32_t), hipMemcpyHostToDevice,
                     s->stream) != hipSuccess) {
    return setError(NBX_ERR_HIP, "target upload failed");
  }
  hipLaunchKernelGGL(k_exact_partial, dim3(blocksFor(nt), kExactChunks),
                     dim3(kBlock), 0, s->stream, static_cast<uint32_t>(s->n), nt,
                     dT, s->px[c], s->py[c], s->m[c], eps2, part);
  hipLaunchKernelGGL(k_exact_reduce, dim3(blocksFor(nt)), dim3(kBlock), 0,
                     s->stream, nt, part, dax, day);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(ax, dax, nt * sizeof(double), hipMemcpyDeviceToHost,
                     s->stream) != hipSuccess ||
      hipMemcpyAsync(ay, day, nt * sizeof(double), hipMemcpyDeviceToHost,
                     s->stream) != hipSuccess ||
      hipStreamSynchronize(s->stream) != hipSuccess) {
    return setError(NBX_ERR_HIP, "exact-force kernels failed");
  }
  return NBX_OK;
}

} // extern "C"

// ---- The learned far-field correction: a network handle (weights, Adam
//      state and a training set in HBM) and its training kernels. ----

struct nbx_mlp {
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  uint64_t cap = 0;   // training pairs the set can hold
  uint64_t count = 0; // pairs held
  float4 *data = nullptr; // 4 float4 per pair: {f0..f3} {f4..f7} {f8 f9 t0 t1} {w}
  float *params = nullptr, *adamM = nullptr, *adamV = nullptr, *grad = nullptr;
  bf16x8 *pk = nullptr; // packed operands (forward + transposed)
  float *bc = nullptr;  // biases, accumulator order
  float *partial = nullptr; // per-workgroup gradient sums
  float *loss = nullptr;    // per-step losses of the last nbx_mlp_train
  uint32_t lossCap = 0;
  uint64_t adamStep = 0;
  bool hasParams = false;
  int4 *pairs = nullptr;    // generation scratch: {body, node, target index}
  double *targetNorm = nullptr; // per target: |monopole sum| of its pairs
  double *wsum = nullptr;   // loss-weight sum
  uint32_t pairCap = 0;
  uint32_t *pairCount = nullptr;
  uint32_t *targets = nullptr;
  uint32_t targetCap = 0;
  uint32_t lastTake = 0;    // pairs of the last generate call (in `pairs`)
  float *io = nullptr;      // staging of predict inputs/outputs
  uint64_t ioCap = 0;
};

namespace {

constexpr uint32_t kMlpGroups = 256;        // workgroups of a training step
constexpr uint32_t kMlpLossCap = 1u << 20;  // steps per nbx_mlp_train call
constexpr uint32_t kMlpPairScratch = 1u << 24;
constexpr uint32_t kMlpTargetScratch = 1u << 22;

template <class T>
int mlpAlloc(nbx_mlp *m, T **p, uint64_t count) {
  void *q = nullptr;
  const size_t bytes = std::max<uint64_t>(count, 1) * sizeof(T);
  if (hipMalloc(&q, bytes) != hipSuccess) {
    return setError(NBX_ERR_HIP, "hipMalloc of " + std::to_string(bytes) +
                                     " bytes failed");
  }
  m->allocs.push_back(q);
  if (hipMemset(q, 0, bytes) != hipSuccess) {
    return setError(NBX_ERR_HIP, "hipMemset failed");
  }
  *p = static_cast<T *>(q);
  return NBX_OK;
}

int mlpPack(nbx_mlp *m) {
  hipLaunchKernelGGL(k_mlp_pack, dim3(blocksFor(kPkAll * 8)), dim3(kBlock), 0,
                     m->stream, m->params,
                     reinterpret_cast<__bf16 *>(m->pk), m->bc);
  return checkLaunch();
}

uint64_t gcd64(uint64_t a, uint64_t b) {
  while (b) {
    const uint64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

// One gradient pass over `batch` positions of the permutation (mulA, addC)
// starting at k0, summed into m->partial; wmean normalises the loss weights.
int mlpGradLaunch(nbx_mlp *m, uint64_t mulA, uint64_t addC, uint64_t k0,
                  uint32_t batch, double wmean) {
  const uint32_t per =
      ((batch + kMlpGroups - 1) / kMlpGroups + kMlpTile - 1) / kMlpTile * kMlpTile;
  const uint32_t groups = (batch + per - 1) / per;
  const float invBW = static_cast<float>(1.0 / (static_cast<double>(batch) * wmean));
  hipLaunchKernelGGL(k_mlp_grad, dim3(groups), dim3(kMlpTrainBlock), 0,
                     m->stream, m->pk, m->bc, m->data, m->count, mulA, addC,
                     k0, batch, per, invBW, m->partial);
  return checkLaunch();
}

// Mean loss weight of pairs [first, first + count) (a zero mean counts as 1).
int mlpWeightMean(nbx_mlp *m, uint64_t first, uint64_t count, double *mean) {
  NBX_HIP(hipMemsetAsync(m->wsum, 0, sizeof(double), m->stream));
  const uint32_t grid = static_cast<uint32_t>(
      std::min<uint64_t>(1024, std::max<uint64_t>(1, blocksFor(count))));
  hipLaunchKernelGGL(k_mlp_wsum, dim3(grid), dim3(kBlock), 0, m->stream, m->data,
                     first, count, m->wsum);
  NBX_TRY(checkLaunch());
  double sum = 0.0;
  NBX_HIP(hipMemcpyAsync(&sum, m->wsum, sizeof(double), hipMemcpyDeviceToHost,
                         m->stream));
  NBX_HIP(hipStreamSynchronize(m->stream));
  *mean = sum > 0.0 ? sum / static_cast<double>(count) : 1.0;
  return NBX_OK;
}

uint32_t mlpGroups(uint32_t batch) {
  const uint32_t per =
      ((batch + kMlpGroups - 1) / kMlpGroups + kMlpTile - 1) / kMlpTile * kMlpTile;
  return (batch + per - 1) / per;
}

} // namespace

extern "C" {

int nbx_mlp_create(uint64_t pair_capacity, nbx_mlp **out) {
  if (!out) {
    return setError(NBX_ERR_ARG, "out is null");
  }
  *out = nullptr;
  if (pair_capacity == 0 || pair_capacity > (1ull << 31)) {
    return setError(NBX_ERR_ARG, "pair capacity must be in [1, 2^31]");
  }
  nbx_mlp *m = new nbx_mlp();
  auto fail = [&](int rc) {
    nbx_mlp_destroy(m);
    return rc;
  };
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
    delete m;
    return setError(NBX_ERR_HIP, "hipStreamCreateWithFlags failed");
  }
  m->cap = pair_capacity;
  m->pairCap = kMlpPairScratch;
  m->targetCap = kMlpTargetScratch;
  m->lossCap = kMlpLossCap;
  int rc = NBX_OK;
  if ((rc = mlpAlloc(m, &m->data, 4 * pair_capacity)) ||
      (rc = mlpAlloc(m, &m->params, kMlpParams)) ||
      (rc = mlpAlloc(m, &m->adamM, kMlpParams)) ||
      (rc = mlpAlloc(m, &m->adamV, kMlpParams)) ||
      (rc = mlpAlloc(m, &m->grad, kMlpParams)) ||
      (rc = mlpAlloc(m, &m->pk, kPkAll)) || (rc = mlpAlloc(m, &m->bc, kBcAll)) ||
      (rc = mlpAlloc(m, &m->partial,
                     static_cast<uint64_t>(kMlpGroups) * (kMlpParams + 1))) ||
      (rc = mlpAlloc(m, &m->loss, m->lossCap)) ||
      (rc = mlpAlloc(m, &m->pairs, m->pairCap)) ||
      (rc = mlpAlloc(m, &m->pairCount, 1)) ||
      (rc = mlpAlloc(m, &m->targets, m->targetCap)) ||
      (rc = mlpAlloc(m, &m->targetNorm, m->targetCap)) ||
      (rc = mlpAlloc(m, &m->wsum, 1))) {
    return fail(rc);
  }
  *out = m;
  return NBX_OK;
}

int nbx_mlp_destroy(nbx_mlp *m) {
  if (!m) {
    return NBX_OK;
  }
  if (m->stream) {
    (void)hipStreamSynchronize(m->stream);
  }
  for (void *p : m->allocs) {
    (void)hipFree(p);
  }
  if (m->stream) {
    (void)hipStreamDestroy(m->stream);
  }
  delete m;
  return NBX_OK;
}

int nbx_mlp_set_params(nbx_mlp *m, const float *params) {
  if (!m || !params) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  for (int i = 0; i < kMlpParams; ++i) {
    if (!std::isfinite(params[i])) {
      return setError(NBX_ERR_ARG, "parameters must be finite");
    }
  }
  NBX_HIP(hipMemcpyAsync(m->params, params, kMlpParams * sizeof(float),
                         hipMemcpyHostToDevice, m->stream));
  NBX_HIP(hipMemsetAsync(m->adamM, 0, kMlpParams * sizeof(float), m->stream));
  NBX_HIP(hipMemsetAsync(m->adamV, 0, kMlpParams * sizeof(float), m->stream));
  m->adamStep = 0;
  NBX_TRY(mlpPack(m));
  NBX_HIP(hipStreamSynchronize(m->stream));
  m->hasParams = true;
  return NBX_OK;
}

int nbx_mlp_get_params(nbx_mlp *m, float *params) {
  if (!m || !params) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  NBX_HIP(hipMemcpyAsync(params, m->params, kMlpParams * sizeof(float),
                         hipMemcpyDeviceToHost, m->stream));
  NBX_HIP(hipStreamSynchronize(m->stream));
  return NBX_OK;
}

int nbx_mlp_clear(nbx_mlp *m) {
  if (!m) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  m->count = 0;
  return NBX_OK;
}

int nbx_mlp_count(nbx_mlp *m, uint64_t *count) {
  if (!m || !count) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  *count = m->count;
  return NBX_OK;
}

int nbx_mlp_generate(nbx_mlp *m, nbx_sim *s, float theta, uint64_t n_targets,
                     const uint32_t *targets, uint64_t *added) {
  if (!m || !s || (n_targets && !targets)) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  if (added) {
    *added = 0;
  }
  if (!(theta > 0.0f)) {
    return setError(NBX_ERR_ARG, "theta must be > 0");
  }
  if (n_targets > m->targetCap) {
    return setError(NBX_ERR_ARG, "at most " + std::to_string(m->targetCap) +
                                     " targets per call");
  }
  for (uint64_t t = 0; t < n_targets; ++t) {
    if (targets[t] >= s->n) {
      return setError(NBX_ERR_ARG, "target index out of range");
    }
  }
  if (!s->treeBuilt || !s->momentsBuilt) {
    return setError(NBX_ERR_STATE,
                    "training pairs need a tree built with node moments (law "
                    "NBX_LAW_ACCURATE_QUAD or NBX_LAW_LEARNED)");
  }
  if (n_targets == 0) {
    return NBX_OK;
  }
  NBX_TRY(syncAndCheck(s));
  const uint32_t nt = static_cast<uint32_t>(n_targets);
  NBX_HIP(hipMemcpyAsync(m->targets, targets, nt * sizeof(uint32_t),
                         hipMemcpyHostToDevice, s->stream));
  NBX_HIP(hipMemsetAsync(m->pairCount, 0, sizeof(uint32_t), s->stream));
  if (!s->learnBuilt) {
    hipLaunchKernelGGL(k_node_learn, dim3(blocksFor(s->nodeCap)), dim3(kBlock), 0,
                       s->stream, s->nodeF, s->nodeQ, s->ctl, s->nodeL);
    NBX_TRY(checkLaunch());
  }
  const float eSq = s->params.epsilon * s->params.epsilon;
  hipLaunchKernelGGL(k_mlp_gen_pairs, dim3(blocksFor(nt)), dim3(kBlock), 0,
                     s->stream, nt, m->targets, s->bodyPack, s->nodeF, s->nodeI,
                     s->centerX, s->centerY, s->size, s->ctl, theta * theta,
                     static_cast<double>(eSq), m->pairs, m->pairCap, m->pairCount,
                     m->targetNorm);
  NBX_TRY(checkLaunch());
  uint32_t found = 0;
  NBX_HIP(hipMemcpyAsync(&found, m->pairCount, sizeof(uint32_t),
                         hipMemcpyDeviceToHost, s->stream));
  NBX_HIP(hipStreamSynchronize(s->stream));
  const uint64_t room = m->cap - m->count;
  const uint32_t take = static_cast<uint32_t>(
      std::min<uint64_t>(std::min<uint64_t>(found, m->pairCap), room));
  if (take > 0) {
    hipLaunchKernelGGL(k_mlp_label, dim3(blocksFor(static_cast<uint64_t>(take) * 64)),
                       dim3(kBlock), 0, s->stream, take, m->pairs, m->targetNorm,
                       s->bodyPack, s->nodeF, s->nodeI, s->nodeL, s->size, eSq,
                       static_cast<double>(eSq), m->data + 4 * m->count);
    NBX_TRY(checkLaunch());
    NBX_HIP(hipStreamSynchronize(s->stream));
  }
  m->count += take;
  m->lastTake = take;
  if (added) {
    *added = take;
  }
  if (found > take) {
    return setError(NBX_ERR_CAPACITY,
                    "training pairs beyond the set's capacity were dropped (" +
                        std::to_string(found - take) + ")");
  }
  return NBX_OK;
}

int nbx_mlp_last_pairs(nbx_mlp *m, uint64_t cap, uint32_t *target,
                       int32_t *node, uint64_t *count) {
  if (!m || !count || (cap && (!target || !node))) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  const uint64_t k = std::min<uint64_t>(cap, m->lastTake);
  std::vector<int4> h(k);
  if (k) {
    NBX_HIP(hipMemcpyAsync(h.data(), m->pairs, k * sizeof(int4),
                           hipMemcpyDeviceToHost, m->stream));
    NBX_HIP(hipStreamSynchronize(m->stream));
  }
  for (uint64_t i = 0; i < k; ++i) {
    target[i] = static_cast<uint32_t>(h[i].x);
    node[i] = h[i].y;
  }
  *count = m->lastTake;
  return NBX_OK;
}

int nbx_mlp_upload(nbx_mlp *m, uint64_t count, const float *features,
                   const float *labels, const float *weights) {
  if (!m || (count && (!features || !labels))) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  if (count > m->cap) {
    return setError(NBX_ERR_ARG, "count above the set's capacity");
  }
  std::vector<float4> h(4 * count);
  for (uint64_t k = 0; k < count; ++k) {
    const float *f = features + k * kMlpIn;
    h[4 * k] = make_float4(f[0], f[1], f[2], f[3]);
    h[4 * k + 1] = make_float4(f[4], f[5], f[6], f[7]);
    h[4 * k + 2] = make_float4(f[8], f[9], labels[2 * k], labels[2 * k + 1]);
    h[4 * k + 3] = make_float4(weights ? weights[k] : 1.0f, 0.f, 0.f, 0.f);
  }
  if (count) {
    NBX_HIP(hipMemcpyAsync(m->data, h.data(), h.size() * sizeof(float4),
                           hipMemcpyHostToDevice, m->stream));
    NBX_HIP(hipStreamSynchronize(m->stream));
  }
  m->count = count;
  return NBX_OK;
}

int nbx_mlp_download(nbx_mlp *m, uint64_t first, uint64_t count,
                     float *features, float *labels, float *weights) {
  if (!m || (count && (!features || !labels))) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  if (first > m->count || count > m->count - first) {
    return setError(NBX_ERR_ARG, "range outside the training set");
  }
  std::vector<float4> h(4 * count);
  if (count) {
    NBX_HIP(hipMemcpyAsync(h.data(), m->data + 4 * first, h.size() * sizeof(float4),
                           hipMemcpyDeviceToHost, m->stream));
    NBX_HIP(hipStreamSynchronize(m->stream));
  }
  for (uint64_t k = 0; k < count; ++k) {
    float *f = features + k * kMlpIn;
    const float4 a = h[4 * k], b = h[4 * k + 1], c = h[4 * k + 2];
    if (weights) {
      weights[k] = h[4 * k + 3].x;
    }
    f[0] = a.x; f[1] = a.y; f[2] = a.z; f[3] = a.w;
    f[4] = b.x; f[5] = b.y; f[6] = b.z; f[7] = b.w;
    f[8] = c.x; f[9] = c.y;
    labels[2 * k] = c.z;
    labels[2 * k + 1] = c.w;
  }
  return NBX_OK;
}

int nbx_mlp_gradient(nbx_mlp *m, uint64_t first, uint64_t count, float *grad,
                     float *loss) {
  if (!m || !grad || !loss) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  if (!m->hasParams) {
    return setError(NBX_ERR_STATE, "no parameters set");
  }
  if (count == 0 || count > 0xffffffffull || first > m->count ||
      count > m->count - first) {
    return setError(NBX_ERR_ARG, "range outside the training set");
  }
  const uint32_t batch = static_cast<uint32_t>(count);
  double wmean = 1.0;
  NBX_TRY(mlpWeightMean(m, first, count, &wmean));
  NBX_TRY(mlpGradLaunch(m, 1, first, 0, batch, wmean));
  hipLaunchKernelGGL(k_mlp_adam, dim3(blocksFor(kMlpParams + 1)), dim3(kBlock), 0,
                     m->stream, m->partial, mlpGroups(batch), 0, 0.0f,
                     0.0f, 0.0f, m->params, m->adamM, m->adamV, m->grad, m->loss);
  NBX_TRY(checkLaunch());
  NBX_HIP(hipMemcpyAsync(grad, m->grad, kMlpParams * sizeof(float),
                         hipMemcpyDeviceToHost, m->stream));
  NBX_HIP(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost,
                         m->stream));
  NBX_HIP(hipStreamSynchronize(m->stream));
  return NBX_OK;
}

int nbx_mlp_train(nbx_mlp *m, uint32_t steps, uint32_t batch, float lr,
                  uint64_t seed, float *losses) {
  if (!m) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  if (!m->hasParams) {
    return setError(NBX_ERR_STATE, "no parameters set");
  }
  if (m->count == 0) {
    return setError(NBX_ERR_STATE, "empty training set");
  }
  if (batch == 0 || steps > m->lossCap || !(lr > 0.0f)) {
    return setError(NBX_ERR_ARG, "batch must be > 0, steps <= 2^20, lr > 0");
  }
  // A permutation of the set, k -> (a k + c) mod count, a coprime to count;
  // consecutive steps take consecutive batches of it.
  const uint64_t count = m->count;
  uint64_t a = (seed * 0x9e3779b97f4a7c15ull) % count;
  a = a | 1u;
  while (gcd64(a, count) != 1) {
    a = (a + 2) % count;
    if (a == 0) a = 1;
  }
  const uint64_t c = (seed * 0xbf58476d1ce4e5b9ull + 12345) % count;
  const uint32_t groups = mlpGroups(batch);
  double wmean = 1.0;
  NBX_TRY(mlpWeightMean(m, 0, count, &wmean));
  for (uint32_t t = 0; t < steps; ++t) {
    NBX_TRY(mlpGradLaunch(m, a, c, static_cast<uint64_t>(t) * batch % count, batch,
                          wmean));
    ++m->adamStep;
    const double st = static_cast<double>(m->adamStep);
    const float bc1 = static_cast<float>(1.0 / (1.0 - std::pow(0.9, st)));
    const float bc2 = static_cast<float>(1.0 / (1.0 - std::pow(0.999, st)));
    hipLaunchKernelGGL(k_mlp_adam, dim3(blocksFor(kMlpParams + 1)), dim3(kBlock), 0,
                       m->stream, m->partial, groups, 1, lr, bc1, bc2,
                       m->params, m->adamM, m->adamV, m->grad, m->loss + t);
    NBX_TRY(checkLaunch());
    NBX_TRY(mlpPack(m));
  }
  if (losses && steps) {
    NBX_HIP(hipMemcpyAsync(losses, m->loss, steps * sizeof(float),
                           hipMemcpyDeviceToHost, m->stream));
  }
  NBX_HIP(hipStreamSynchronize(m->stream));
  return NBX_OK;
}

int nbx_mlp_predict(nbx_mlp *m, uint64_t count, const float *features,
                    float *out) {
  if (!m || (count && (!features || !out))) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  if (!m->hasParams) {
    return setError(NBX_ERR_STATE, "no parameters set");
  }
  if (count == 0) {
    return NBX_OK;
  }
  if (count > 0xffffffffull / 16) {
    return setError(NBX_ERR_ARG, "too many pairs");
  }
  const uint64_t need = count * (kMlpIn + kMlpOut);
  if (need > m->ioCap) {
    float *p = nullptr;
    if (hipMalloc(&p, need * sizeof(float)) != hipSuccess) {
      return setError(NBX_ERR_HIP, "hipMalloc failed in nbx_mlp_predict");
    }
    m->allocs.push_back(p);
    m->io = p;
    m->ioCap = need;
  }
  float *fin = m->io, *fout = m->io + count * kMlpIn;
  NBX_HIP(hipMemcpyAsync(fin, features, count * kMlpIn * sizeof(float),
                         hipMemcpyHostToDevice, m->stream));
  const uint64_t waves = (count + kMlpTile - 1) / kMlpTile;
  hipLaunchKernelGGL(k_mlp_predict, dim3(blocksFor(waves * 64)), dim3(kBlock), 0,
                     m->stream, m->pk, m->bc, fin, static_cast<uint32_t>(count),
                     fout);
  NBX_TRY(checkLaunch());
  NBX_HIP(hipMemcpyAsync(out, fout, count * kMlpOut * sizeof(float),
                         hipMemcpyDeviceToHost, m->stream));
  NBX_HIP(hipStreamSynchronize(m->stream));
  return NBX_OK;
}

int nbx_set_mlp(nbx_sim *s, nbx_mlp *m) {
  if (!s || !m) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  if (!m->hasParams) {
    return setError(NBX_ERR_STATE, "no parameters set");
  }
  NBX_HIP(hipStreamSynchronize(m->stream));
  NBX_HIP(hipMemcpyAsync(s->mlpPk, m->pk, kPkFwd * sizeof(bf16x8),
                         hipMemcpyDeviceToDevice, s->stream));
  NBX_HIP(hipMemcpyAsync(s->mlpBc, m->bc, kBcAll * sizeof(float),
                         hipMemcpyDeviceToDevice, s->stream));
  NBX_HIP(hipStreamSynchronize(s->stream));
  s->mlpLoaded = true;
  return NBX_OK;
}

} // extern "C"

// ---- Ensembles: many independent simulations stepped together (two
//      dispatches a step for all of them, captured in a HIP graph). ----

struct nbx_ensemble {
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  std::vector<EnsInst> inst; // host copy (pointers, sizes; `cur` mirrored)
  std::vector<uint64_t> cap;
  EnsInst *dInst = nullptr;
  uint32_t *dOrder = nullptr, *dWaveInst = nullptr, *dQueue = nullptr;
  uint32_t totalWaves = 0;
  nbx_params params{};
  uint32_t prepareGrid = 0;           // workgroups of k_ens_prepare
  uint32_t walkBlock = kEnsWalkBlock; // lanes per workgroup of k_ens_walk
  uint64_t frame = 0;
  hipGraphExec_t graph = nullptr;
  bool instDirty = true; // host table changed: upload before the next step
};

namespace {

template <class T>
int ensAlloc(nbx_ensemble *e, T **p, uint64_t count) {
  void *q = nullptr;
  const size_t bytes = std::max<uint64_t>(count, 1) * sizeof(T);
  if (hipMalloc(&q, bytes) != hipSuccess) {
    return setError(NBX_ERR_HIP, "hipMalloc of " + std::to_string(bytes) +
                                     " bytes failed");
  }
  e->allocs.push_back(q);
  if (hipMemset(q, 0, bytes) != hipSuccess) {
    return setError(NBX_ERR_HIP, "hipMemset failed");
  }
  *p = static_cast<T *>(q);
  return NBX_OK;
}

int ensCheck(nbx_ensemble *e, uint32_t k) {
  if (!e) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  if (k >= e->inst.size()) {
    return setError(NBX_ERR_ARG, "instance index out of range");
  }
  return NBX_OK;
}

void ensDropGraph(nbx_ensemble *e) {
  if (e->graph) {
    (void)hipGraphExecDestroy(e->graph);
    e->graph = nullptr;
  }
}

// Instance order for the persistent kernel (largest first) and the flat
// wave table of the walk.
int ensLayout(nbx_ensemble *e) {
  const uint32_t count = static_cast<uint32_t>(e->inst.size());
  std::vector<uint32_t> order(count);
  for (uint32_t k = 0; k < count; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return e->inst[a].n > e->inst[b].n;
  });
  std::vector<uint32_t> waveInst;
  for (uint32_t k = 0; k < count; ++k) {
    e->inst[k].waveBase = static_cast<uint32_t>(waveInst.size());
    const uint32_t waves = (e->inst[k].n + kWave - 1) / kWave;
    waveInst.insert(waveInst.end(), waves, k);
  }
  e->totalWaves = static_cast<uint32_t>(waveInst.size());
  NBX_HIP(hipMemcpy(e->dOrder, order.data(), count * sizeof(uint32_t),
                    hipMemcpyHostToDevice));
  if (!waveInst.empty()) {
    NBX_HIP(hipMemcpy(e->dWaveInst, waveInst.data(),
                      waveInst.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  e->instDirty = true;
  ensDropGraph(e);
  return NBX_OK;
}

int ensPushTable(nbx_ensemble *e) {
  if (e->instDirty) {
    NBX_HIP(hipMemcpyAsync(e->dInst, e->inst.data(), e->inst.size() * sizeof(EnsInst),
                           hipMemcpyHostToDevice, e->stream));
    NBX_HIP(hipStreamSynchronize(e->stream));
    e->instDirty = false;
  }
  return NBX_OK;
}

// `cur` of every instance as the device has it.
int ensPullCur(nbx_ensemble *e) {
  std::vector<EnsInst> d(e->inst.size());
  NBX_HIP(hipMemcpyAsync(d.data(), e->dInst, d.size() * sizeof(EnsInst),
                         hipMemcpyDeviceToHost, e->stream));
  NBX_HIP(hipStreamSynchronize(e->stream));
  for (size_t k = 0; k < d.size(); ++k) {
    e->inst[k].cur = d[k].cur;
  }
  return NBX_OK;
}

int ensLaunchStep(nbx_ensemble *e) {
  const float tSq = e->params.theta * e->params.theta;
  const float eSq = e->params.epsilon * e->params.epsilon;
  EnsParams p{e->params.dt, tSq, eSq, e->params.collisions};
  hipLaunchKernelGGL(k_ens_reset, dim3(1), dim3(1), 0, e->stream, e->dQueue);
  hipLaunchKernelGGL(k_ens_prepare, dim3(e->prepareGrid), dim3(kEnsBlock), 0,
                     e->stream, e->dInst, e->dOrder,
                     static_cast<uint32_t>(e->inst.size()), p, e->dQueue);
  const bool acc = e->params.law == NBX_LAW_ACCURATE;
  auto walk = [&](auto block) {
    constexpr int kB = decltype(block)::value;
    const uint32_t blocks = (e->totalWaves + kB / kWave - 1) / (kB / kWave);
    if (!blocks) {
      return;
    }
    if (acc) {
      hipLaunchKernelGGL((k_ens_walk<kLawAccurate, kB>), dim3(blocks), dim3(kB), 0,
                         e->stream, e->dInst, e->dWaveInst, e->totalWaves, tSq, eSq);
    } else {
      hipLaunchKernelGGL((k_ens_walk<kLawReference, kB>), dim3(blocks), dim3(kB), 0,
                         e->stream, e->dInst, e->dWaveInst, e->totalWaves, tSq, eSq);
    }
  };
  if (e->walkBlock == 64) {
    walk(std::integral_constant<int, 64>{});
  } else if (e->walkBlock == 256) {
    walk(std::integral_constant<int, 256>{});
  } else {
    walk(std::integral_constant<int, 128>{});
  }
  return checkLaunch();
}

} // namespace

extern "C" {

int nbx_ensemble_create(uint32_t count, const uint64_t *capacity,
                        nbx_ensemble **out) {
  if (!out || (count && !capacity)) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  *out = nullptr;
  if (count == 0 || count > 65536) {
    return setError(NBX_ERR_ARG, "instance count must be in [1, 65536]");
  }
  for (uint32_t k = 0; k < count; ++k) {
    if (capacity[k] == 0 || capacity[k] > (1ull << 20)) {
      return setError(NBX_ERR_ARG, "instance capacity must be in [1, 2^20]");
    }
  }
  nbx_ensemble *e = new nbx_ensemble();
  auto fail = [&](int rc) {
    nbx_ensemble_destroy(e);
    return rc;
  };
  nbx_default_params(&e->params);
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
    delete e;
    return setError(NBX_ERR_HIP, "hipStreamCreateWithFlags failed");
  }
  e->inst.resize(count);
  e->cap.assign(capacity, capacity + count);
  int rc = NBX_OK;
#define EA(ptr, cnt)                                                          \
  if ((rc = ensAlloc(e, &(ptr), (cnt))) != NBX_OK)                            \
    return fail(rc);
  for (uint32_t k = 0; k < count; ++k) {
    EnsInst &I = e->inst[k];
    const uint64_t n = capacity[k];
    I.n = 0;
    I.cur = 0;
    I.buckets = static_cast<uint32_t>(std::max<uint64_t>(1024, pow2AtLeast(2 * n)));
    I.pairCap = static_cast<uint32_t>(8 * n + 1024);
    I.nodeCap = static_cast<int32_t>(8 * n + 8);
    I.cell = I.invCell = 1.0f;
    for (int b = 0; b < 2; ++b) {
      EA(I.px[b], n) EA(I.py[b], n) EA(I.vx[b], n) EA(I.vy[b], n)
      EA(I.m[b], n) EA(I.r[b], n) EA(I.id[b], n)
      EA(I.keyHi[b], n) EA(I.keyLo[b], n) EA(I.val[b], n)
    }
    EA(I.ax, n) EA(I.ay, n) EA(I.body, n + kBodyPad)
    EA(I.bucketCount, I.buckets + 1) EA(I.bucketRank, n) EA(I.bucketOf, n)
    I.bucketStart = I.bucketCount;
    EA(I.grid, 2 * n) EA(I.largeList, n) EA(I.largeBox, n)
    EA(I.pairs, I.pairCap) EA(I.seg, I.pairCap) EA(I.parent, n)
    EA(I.compCount, n + 1) EA(I.compStart, n + 1)
    EA(I.nodeF, I.nodeCap) EA(I.nodeI, I.nodeCap) EA(I.centerX, I.nodeCap)
    EA(I.centerY, I.nodeCap) EA(I.size, I.nodeCap) EA(I.wsum, I.nodeCap)
    EA(I.ctl, 1)
  }
  EA(e->dInst, count) EA(e->dOrder, count) EA(e->dQueue, 1)
  {
    uint64_t waves = 0;
    for (uint32_t k = 0; k < count; ++k) waves += (capacity[k] + kWave - 1) / kWave;
    EA(e->dWaveInst, waves)
  }
#undef EA
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
          hipSuccess) {
    return fail(setError(NBX_ERR_HIP, "device query failed"));
  }
  e->prepareGrid = std::min<uint32_t>(count, 2u * static_cast<uint32_t>(cus));
  if ((rc = ensLayout(e)) != NBX_OK) {
    return fail(rc);
  }
  *out = e;
  return NBX_OK;
}

int nbx_ensemble_destroy(nbx_ensemble *e) {
  if (!e) {
    return NBX_OK;
  }
  if (e->stream) {
    (void)hipStreamSynchronize(e->stream);
  }
  ensDropGraph(e);
  for (void *p : e->allocs) {
    (void)hipFree(p);
  }
  if (e->stream) {
    (void)hipStreamDestroy(e->stream);
  }
  delete e;
  return NBX_OK;
}

int nbx_ensemble_upload(nbx_ensemble *e, uint32_t k, uint64_t n, const float *px,
                        const float *py, const float *vx, const float *vy,
                        const float *ax, const float *ay, const float *m,
                        const float *r, const uint32_t *id) {
  NBX_TRY(ensCheck(e, k));
  if (n && (!px || !py || !vx || !vy || !m || !r || !id)) {
    return setError(NBX_ERR_ARG, "null array");
  }
  if (n > e->cap[k]) {
    return setError(NBX_ERR_ARG, "n exceeds the instance's capacity");
  }
  NBX_HIP(hipStreamSynchronize(e->stream));
  NBX_TRY(ensPullCur(e));
  EnsInst &I = e->inst[k];
  I.n = static_cast<uint32_t>(n);
  I.cur = 0;
  // The grid cell: 1.01 x the median box width (Collision.cpp:173-183),
  // fixed by the radii, which never change. Any positive cell gives the
  // same pair set; this one keeps the stencil small.
  float cell = 1.0f;
  if (n > 0) {
    std::vector<float> radii(r, r + n);
    std::nth_element(radii.begin(), radii.begin() + n / 2, radii.end());
    cell = 2.0f * radii[n / 2] * 1.01f;
  }
  if (!(cell > 0.0f) || !std::isfinite(cell)) {
    cell = 1.0f;
  }
  I.cell = cell;
  I.invCell = 1.0f / cell;
  const size_t fb = n * sizeof(float);
  NBX_HIP(hipMemcpy(I.px[0], px, fb, hipMemcpyHostToDevice));
  NBX_HIP(hipMemcpy(I.py[0], py, fb, hipMemcpyHostToDevice));
  NBX_HIP(hipMemcpy(I.vx[0], vx, fb, hipMemcpyHostToDevice));
  NBX_HIP(hipMemcpy(I.vy[0], vy, fb, hipMemcpyHostToDevice));
  NBX_HIP(hipMemcpy(I.m[0], m, fb, hipMemcpyHostToDevice));
  NBX_HIP(hipMemcpy(I.r[0], r, fb, hipMemcpyHostToDevice));
  NBX_HIP(hipMemcpy(I.id[0], id, n * sizeof(uint32_t), hipMemcpyHostToDevice));
  if (ax) {
    NBX_HIP(hipMemcpy(I.ax, ax, fb, hipMemcpyHostToDevice));
  } else {
    NBX_HIP(hipMemset(I.ax, 0, fb));
  }
  if (ay) {
    NBX_HIP(hipMemcpy(I.ay, ay, fb, hipMemcpyHostToDevice));
  } else {
    NBX_HIP(hipMemset(I.ay, 0, fb));
  }
  NBX_HIP(hipMemset(I.ctl, 0, sizeof(Ctl)));
  return ensLayout(e);
}

int nbx_ensemble_init(nbx_ensemble *e, uint32_t k, int scenario, uint64_t n,
                      uint32_t seed) {
  NBX_TRY(ensCheck(e, k));
  if (n > e->cap[k]) {
    return setError(NBX_ERR_ARG, "n exceeds the instance's capacity");
  }
  std::vector<float> px(n), py(n), vx(n), vy(n), m(n), r(n);
  std::vector<uint32_t> id(n);
  switch (scenario) {
  case NBX_SCENARIO_DISC:
    nbx_host::uniformDisc(n, seed, px.data(), py.data(), vx.data(), vy.data(),
                          m.data(), r.data(), id.data());
    break;
  case NBX_SCENARIO_TWO_GALAXIES:
  case NBX_SCENARIO_MERGER:
    if (n < 2) {
      return setError(NBX_ERR_ARG, "the two-disc scenarios need n >= 2");
    }
    (scenario == NBX_SCENARIO_MERGER ? nbx_host::galaxyMerger : nbx_host::twoGalaxies)(
        n, seed, px.data(), py.data(), vx.data(), vy.data(), m.data(), r.data(),
        id.data());
    break;
  case NBX_SCENARIO_COLD_SQUARE:
    nbx_host::coldSquare(n, seed, px.data(), py.data(), vx.data(), vy.data(),
                         m.data(), r.data(), id.data());
    break;
  default:
    return setError(NBX_ERR_ARG, "unknown scenario");
  }
  return nbx_ensemble_upload(e, k, n, px.data(), py.data(), vx.data(), vy.data(),
                             nullptr, nullptr, m.data(), r.data(), id.data());
}

int nbx_ensemble_download(nbx_ensemble *e, uint32_t k, float *px, float *py,
                          float *vx, float *vy, float *ax, float *ay, float *m,
                          float *r, uint32_t *id) {
  NBX_TRY(ensCheck(e, k));
  NBX_HIP(hipStreamSynchronize(e->stream));
  NBX_TRY(ensPullCur(e));
  const EnsInst &I = e->inst[k];
  const int c = static_cast<int>(I.cur);
  const size_t fb = I.n * sizeof(float);
  if (I.n == 0) {
    return NBX_OK;
  }
  if (px) NBX_HIP(hipMemcpy(px, I.px[c], fb, hipMemcpyDeviceToHost));
  if (py) NBX_HIP(hipMemcpy(py, I.py[c], fb, hipMemcpyDeviceToHost));
  if (vx) NBX_HIP(hipMemcpy(vx, I.vx[c], fb, hipMemcpyDeviceToHost));
  if (vy) NBX_HIP(hipMemcpy(vy, I.vy[c], fb, hipMemcpyDeviceToHost));
  if (ax) NBX_HIP(hipMemcpy(ax, I.ax, fb, hipMemcpyDeviceToHost));
  if (ay) NBX_HIP(hipMemcpy(ay, I.ay, fb, hipMemcpyDeviceToHost));
  if (m) NBX_HIP(hipMemcpy(m, I.m[c], fb, hipMemcpyDeviceToHost));
  if (r) NBX_HIP(hipMemcpy(r, I.r[c], fb, hipMemcpyDeviceToHost));
  if (id) NBX_HIP(hipMemcpy(id, I.id[c], I.n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return NBX_OK;
}

int nbx_ensemble_set_params(nbx_ensemble *e, const nbx_params *p) {
  if (!e || !p) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  NBX_TRY(validParams(p));
  if (p->law != NBX_LAW_REFERENCE && p->law != NBX_LAW_ACCURATE) {
    return setError(NBX_ERR_ARG, "ensembles support the reference and accurate laws");
  }
  if (p->summation != NBX_SUM_COMPENSATED ||
      p->collide_mode != NBX_COLLIDE_SEQUENTIAL) {
    return setError(NBX_ERR_ARG,
                    "ensembles run compensated summation and the sequential collide");
  }
  if (std::memcmp(&e->params, p, sizeof *p) != 0) {
    NBX_HIP(hipStreamSynchronize(e->stream));
    ensDropGraph(e);
    e->params = *p;
  }
  return NBX_OK;
}

int nbx_ensemble_set_launch(nbx_ensemble *e, uint32_t walkLanes,
                            uint32_t prepareGroups) {
  if (!e) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  if (walkLanes != 0 && walkLanes != 64 && walkLanes != 128 && walkLanes != 256) {
    return setError(NBX_ERR_ARG, "walk workgroup size must be 64, 128 or 256");
  }
  if (prepareGroups > e->inst.size()) {
    return setError(NBX_ERR_ARG, "more per-instance workgroups than instances");
  }
  const uint32_t wl = walkLanes ? walkLanes : e->walkBlock;
  const uint32_t pg = prepareGroups ? prepareGroups : e->prepareGrid;
  if (wl != e->walkBlock || pg != e->prepareGrid) {
    NBX_HIP(hipStreamSynchronize(e->stream));
    ensDropGraph(e);
    e->walkBlock = wl;
    e->prepareGrid = pg;
  }
  return NBX_OK;
}

int nbx_ensemble_get_launch(nbx_ensemble *e, uint32_t *walkLanes,
                            uint32_t *prepareGroups) {
  if (!e || !walkLanes || !prepareGroups) {
    return setError(NBX_ERR_ARG, "null argument");
  }
  *walkLanes = e->walkBlock;
  *prepareGroups = e->prepareGrid;
  return NBX_OK;
}

int nbx_ensemble_step(nbx_ensemble *e, uint32_t steps) {
  if (!e) {
    return setError(NBX_ERR_ARG, "null handle");
  }
  NBX_TRY(ensPushTable(e));
  if (e->params.use_graph && !e->graph && steps > 0) {
    hipGraph_t g = nullptr;
    NBX_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    const int rc = ensLaunchStep(e);
    hipError_t ec = hipStreamEndCapture(e->stream, &g);
    if (rc != NBX_OK || ec != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return rc != NBX_OK ? rc : setError(NBX_ERR_HIP, "graph capture failed");
    }
    ec = hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ec != hipSuccess) {
      e->graph = nullptr;
      return setError(NBX_ERR_HIP, "graph instantiation failed");
    }
  }
  for (uint32_t s = 0; s < steps; ++s) {
    if (e->graph) {
      NBX_HIP(hipGraphLaunch(e->graph, e->stream));
    } else {
      NBX_TRY(ensLaunchStep(e));
    }
    ++e->frame;
  }
  NBX_HIP(hipStreamSynchronize(e->stream));
  // Device-side failures (capacity) of any instance.
  std::vector<Ctl> ctl(e->inst.size());
  for (size_t k = 0; k < e->inst.size(); ++k) {
    NBX_HIP(hipMemcpyAsync(&ctl[k], e->inst[k].ctl, sizeof(Ctl), hipMemcpyDeviceToHost,
                           e->stream));
  }
  NBX_HIP(hipStreamSynchronize(e->stream));
  for (size_t k = 0; k < ctl.size(); ++k) {
    if (ctl[k].status & (kStatusNodeOverflow | kStatusPairOverflow)) {
      return setError(NBX_ERR_CAPACITY, "instance " + std::to_string(k) +
                                            " exceeded its node or pair capacity");
    }
    if (ctl[k].status & kStatusTreeInconsistent) {
      return setError(NBX_ERR_STATE, "instance " + std::to_string(k) +
                                         ": inconsistent tree");
    }
  }
  return NBX_OK;
}

} // extern "C"
