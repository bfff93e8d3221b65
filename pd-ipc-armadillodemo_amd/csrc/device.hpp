// Internal: device-side data of one simulator handle, and the launch wrappers
// each kernel file provides.  Positions and every 3-vector field are stored SoA
// with leading dimension `ld` (component c of point i at [c * ld + i]); the
// host boundary converts from / to the reference's AoS xyz layout.
#pragma once
#include <hip/hip_runtime.h>

#include <hip/hip_ext.h>

#include <cstdint>
#include <vector>

namespace pdipc {

constexpr int kWave = 64;

// one row (x, y, z, unused) of an FP64 work vector of the Chebyshev sweeps
struct __align__(32) Row64 {
  double x, y, z, w;
};

// LBVH node for the stackless traversal: box rounded outward to float (a
// superset of the exact box, used only to prune), the left child (internal
// nodes) or the patch id (leaves), and the next node in depth-first preorder
// once this subtree is done (-1: traversal over).
struct __align__(16) BvhNode {
  float lo[3];
  int child;
  float hi[3];
  int skip;
};

// In-step kernel timing (pdipc_kernel_timing): while a timer is attached to
// the mesh, every launch of the timed kernels records its own start and stop
// with HIP events (hipExtLaunchKernel), i.e. the kernel's execution time as a
// profiler reports it; the broad phase is timed as one span.
enum TimerClass {
  TC_SWEEP64 = 0,  // FP64 Chebyshev sweeps
  TC_SWEEP32,      // FP32-split sweeps
  TC_RHS,          // local step: the rhs gather (k_rhs)
  TC_PROJECT,      // local step: cloth curvature and tet projection
  TC_FIRST,        // first FP64 sweep, FP32-split residual passes
  TC_OPERATORS,    // M^2 operator update
  TC_CCD,          // time of impact
  TC_BARRIER,      // barrier projection terms
  TC_BROAD,        // broad phase (span, host waits included)
  TC_COUNT
};

struct KernelTimer {
  std::vector<hipEvent_t> pool;  // (start, stop) pairs
  std::vector<int> cls;          // class of each pair in use
  size_t used = 0;
  double us[TC_COUNT] = {};
  long long count[TC_COUNT] = {};
  // a (start, stop) pair for one launch of class c (false: no event available)
  bool take(int c, hipEvent_t* e0, hipEvent_t* e1) {
    if (2 * used + 2 > pool.size()) {
      hipEvent_t a = nullptr, b = nullptr;
      if (hipEventCreate(&a) != hipSuccess) return false;
      if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return false;
      }
      pool.push_back(a);
      pool.push_back(b);
    }
    *e0 = pool[2 * used];
    *e1 = pool[2 * used + 1];
    if (cls.size() <= used) cls.resize(used + 1);
    cls[used++] = c;
    return true;
  }
  ~KernelTimer() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  }
};

// a kernel launch, timed when a timer is attached
template <class... KA, class... A>
inline void launch_t(KernelTimer* t, int cls, void (*kernel)(KA...), dim3 grid, dim3 block, uint32_t shm,
                     hipStream_t s, A... args) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (t && t->take(cls, &e0, &e1))
    hipExtLaunchKernelGGL(kernel, grid, block, shm, s, e0, e1, 0, args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, shm, s, args...);
}

struct DeviceMesh {
  KernelTimer* timer = nullptr;  // in-step kernel timing (null: off)
  int nTotal = 0, nSim = 0, nCloth = 0, nTets = 0;
  int ld = 0;  // leading dimension of SoA vectors (>= nTotal, multiple of 64)
  // state and work vectors, SoA [3][ld] doubles
  double *X = nullptr, *Xold = nullptr, *X0 = nullptr, *Dir = nullptr, *Pre = nullptr, *Next = nullptr, *V = nullptr;
  double *initRhs = nullptr, *sysRhs = nullptr, *firstTerm = nullptr, *colRhs = nullptr, *clothA = nullptr, *Xrest = nullptr;
  // FP64 sweep work vectors, one Row64 per row (ld rows): three iterates and c
  Row64 *xa0 = nullptr, *xa1 = nullptr, *xa2 = nullptr, *cA = nullptr;
  // FP32-split work vectors, one float4 (x, y, z, 0) per row (ld rows): a
  // sweep gathers one 16-byte vector per matrix entry
  float4 *dPre = nullptr, *dCur = nullptr, *dNext = nullptr, *r0 = nullptr;
  // per point
  double *mass = nullptr, *invMass = nullptr, *gravY = nullptr, *fixK = nullptr, *colMass = nullptr;
  int* fixFlag = nullptr;
  double *inertialDiag = nullptr, *elasticDiag = nullptr, *sysDiag = nullptr, *colDiag = nullptr;
  // rows whose diagonal changed since the last M^2 update; dStale forces rows
  // (all, after the host wrote sysDiag directly); dirty = changed rows and
  // their R1 neighbours (the rows whose M^2 entries depend on a changed D)
  uint8_t *dStale = nullptr, *dChanged = nullptr, *dDirty = nullptr;
  // cloth
  int *cCenter = nullptr, *cOff = nullptr, *cNbr = nullptr, *clothOfPoint = nullptr;
  int2* ringRange = nullptr;  // per simulated point: its cloth ring slots [x, y) (empty for tet points)
  double *cRest = nullptr, *cLbo = nullptr, *cBendIn = nullptr, *cStrain = nullptr, *cH0 = nullptr;
  // tets
  int4* tetIdx = nullptr;
  double* tetDminv = nullptr;  // SoA [9][T]
  double* tetW = nullptr;
  double* tetOut = nullptr;    // SoA [12][T]
  int *vtOff = nullptr, *vtList = nullptr;
  // R1 (off-diagonal CSR)
  int *r1Off = nullptr, *r1Col = nullptr;
  double* r1Val = nullptr;
  // M^2 in SELL layout
  int sellLanes = 4, rowsPerSlice = 16, nSlices = 0;
  // > 0: every slice has this width (offsets by arithmetic); kSellLaneMajorW:
  // the slices are lane-major, otherwise chunk-major
  int sellUniformW = 0;
  int* sellPerm = nullptr;    // SELL slot -> row (nSlices * rowsPerSlice); null: identity
  int* sellSlotOf = nullptr;  // row -> SELL slot (nSim); null: identity
  int64_t sellEntries = 0;
  int64_t* sliceOff = nullptr;
  int* sliceW = nullptr;
  int *m2Col = nullptr, *m2RowFlag = nullptr;
  int16_t* m2Col16 = nullptr;  // col - row per entry, or null when an offset exceeds 16 bits
  double *m2R1 = nullptr, *m2Val = nullptr;
  float* m2ValF = nullptr;
  int64_t* p2Off = nullptr;
  int* p2K = nullptr;
  double* p2A = nullptr;
  // LDS staging of the sweeps' operand (null: the sweeps gather from L2): per
  // tile of kTileSlices slices the union of its columns (tileBase[t] +
  // tileCol[t * stageStride + u], each list padded to stageStride entries with
  // its last column) and per entry its index in that union
  int* tileBase = nullptr;
  uint16_t *tileCol = nullptr, *m2Loc = nullptr;
  int stageStride = 0;
  // narrow staging of the FP32 and FP64 sweeps (lane-major layouts at 8 lanes
  // per row): tiles of kNarrowTileSlices slices, lists at narrowStride,
  // one-byte union indices per entry; null when not used
  int* ntileBase = nullptr;
  uint16_t* ntileCol = nullptr;
  uint8_t* m2Loc8 = nullptr;
  int narrowStride = 0;
  // collision topology
  int nFaces = 0, nEdges = 0, nPatches = 0;
  int3* bFaces = nullptr;
  int2* bEdges = nullptr;
  int *pfOff = nullptr, *pfList = nullptr, *ppOff = nullptr, *ppList = nullptr, *peOff = nullptr, *peList = nullptr;
  uint8_t* patchStatic = nullptr;
  // broad phase work
  double* patchBox = nullptr;      // SoA [6][P]: min xyz, max xyz
  double* sceneBox = nullptr;      // 6 (centroid bounds)
  double* sceneBoxPart = nullptr;  // 6 * parts
  uint32_t *mortonA = nullptr, *mortonB = nullptr;
  int *idxA = nullptr, *idxB = nullptr;
  uint64_t* keys = nullptr;        // sorted (code << 32 | patch)
  int *radixHist = nullptr, *radixBlockSums = nullptr;
  int *nodeLeft = nullptr, *nodeRight = nullptr, *nodeParent = nullptr;
  int2* nodeRange = nullptr;       // leaf range [first, last] of each internal node
  double* nodeTable = nullptr;     // sparse table of leaf-box unions, levels x [6][P]
  BvhNode* nodes = nullptr;        // 2P-1: internal nodes 0..P-2, leaf i (sorted position) at P-1+i
  int2* pairs = nullptr;
  int pairCap = 0;
  // traversal output (lengths cnt*, offsets off*): pass 1 work items
  // (patch, subtree root) and pairs in one segment per wavefront, pass 2
  // pairs in one segment per work item (itemCap segments)
  int2 *segI = nullptr, *segP = nullptr, *segS = nullptr;
  int segCapI = 0, segCapP = 0;    // entries per pass-1 segment
  int *cntI = nullptr, *offI = nullptr, *cntP = nullptr, *offP = nullptr, *cntS = nullptr, *offS = nullptr;
  int2* items = nullptr;           // compacted work items
  int itemCap = 0;
  // 2 * pairCap each: point-triangle counts / offsets per pair, then edge-edge ones
  int *recCount = nullptr, *recOff = nullptr;
  int4* recs = nullptr;            // EE records store ~a in .x
  int recCap = 0;
  int* scanBlock = nullptr;        // scan scratch
  int scanBlockCap = 0;
  // collision terms, summed per point in record order
  int actCap = 0;                  // capacity of the active-record list
  int* actRec = nullptr;           // active record ids
  double *actW = nullptr, *actT = nullptr;  // weight, 4 targets x 3 per active record
  int *ptCount = nullptr, *ptOff = nullptr, *ptCursor = nullptr, *ptScan = nullptr;  // nSim + 1
  unsigned long long* entKey = nullptr;  // 4 * actCap: (record << 2 | corner)
  int *entA = nullptr, *entPt = nullptr;  // 4 * actCap: active-list index, point
  unsigned long long* entSorted = nullptr;  // 4 * actCap: (active index << 2 | corner), key order per point
  // scalars
  unsigned long long* toiBits = nullptr;
  double* partials = nullptr;      // reduction partials
  double* residual = nullptr;
  // [0] pass-2 pairs, [1] records, [2] active records, [3] pass-1 segment
  // overflow (largest length needed), [4] TOI argmin (recorded steps),
  // [5] active-list overflow, [6] work items, [7] pass-1 pairs, [8] edge-edge
  // records ([1]: point-triangle records)
  int* counters = nullptr;
  int nPartials = 0;
  unsigned long long* checkCount = nullptr;  // the zero-intersection check's counter
};

// ---- kernel launch wrappers (all asynchronous on `s`) -------------------------
// k_step.hip
void launch_predict(const DeviceMesh& m, double dt, hipStream_t s);
void launch_diag(const DeviceMesh& m, int withCollision, hipStream_t s);
void launch_direction(const DeviceMesh& m, hipStream_t s);  // also sets the TOI minimum to 1
void launch_clamp_residual(const DeviceMesh& m, double toiScale, hipStream_t s);
void launch_velocity(const DeviceMesh& m, double dt, hipStream_t s);
void launch_zero(double* p, int64_t n, hipStream_t s);
void launch_copy(double* dst, const double* src, int64_t n, hipStream_t s);
// k_local.hip
// residualOmega >= 0 (FP32-split): also the first pass of the solve's FP64
// residual (into xa0 / cA, for the current X and D); the solve is then
// launched with residualReady
void launch_local(const DeviceMesh& m, int precision, double sigLo, double sigHi, int withCollision, hipStream_t s,
                  double residualOmega = -1.0);
// k_solve.hip
void launch_operators(const DeviceMesh& m, double omega, int precision, hipStream_t s);
void launch_solve(const DeviceMesh& m, double omega, double rho, int chebIters, int precision, hipStream_t s,
                  bool residualReady = false);
// one Chebyshev sweep of the given precision (measurement support)
void launch_one_sweep(const DeviceMesh& m, int precision, hipStream_t s);
// k_contact.hip
// ---- adjoint (reverse mode of the FP64 step; active sets frozen)
// index maps and transposed weights uploaded on the first adjoint call
struct AdjMesh {
  int* sellR1Idx = nullptr;  // per SELL entry: R1 index of (i,j) or -1
  int *pIa = nullptr, *pIb = nullptr;  // per 2-path (SELL order): R1 indices of its factors
  int* cSlotR1 = nullptr;    // per cloth slot: R1 index of (center, nbr) or -1
  double *cBendOut = nullptr, *cLboIn = nullptr;  // per cloth slot
  // the 2-paths grouped by their middle point k: (SELL entry, A_p)
  int* midOff = nullptr;
  int64_t* midEnt = nullptr;
  double* midA = nullptr;
  // per spring R1 entry q: the 2-paths using q as a factor (SELL entry, middle point, other factor)
  int nSpringQ = 0;
  int *springQ = nullptr, *springOff = nullptr, *springMid = nullptr, *springOther = nullptr;
  int64_t* springEnt = nullptr;
  // per slice: the sorted union of the columns of its entries (uOff, uCol), and
  // per union column and row of the slice the entry's offset in the slice's
  // SELL block (uPos, 0xFFFF where the row has no entry in that column)
  int *uOff = nullptr, *uCol = nullptr;
  uint16_t* uPos = nullptr;
};
void launch_solve_taped(const DeviceMesh& m, double omega, double rho, int iters, double* const* outs, hipStream_t s);
// the whole reverse of one K-sweep solve: u[k] for k = K-1 .. 0 from u[K] (given)
// and u[K+1] (zero), then one SELL pass adding sum_k W_k u_{k+1} x_k^T to Gbar,
// with the forward iterates x_0 and x_k = x1 + (k-1) xStride (k >= 1)
constexpr int kMaxCheb = 32;
void launch_adj_pass(const DeviceMesh& m, const AdjMesh& a, const double* W, int K, double* const* u,
                     const double* x0, const double* x1, int64_t xStride, double* Gbar, double* uc, hipStream_t s);
void launch_adj_first(const DeviceMesh& m, double omega, const double* b, const double* uc, double* bbar, double* Dbar,
                      double* R1bar, hipStream_t s);
// adjoint of the operator update (M^2 from D and R1) as three deterministic
// gathers over the lists in AdjMesh
void launch_adj_ops_gather(const DeviceMesh& m, const AdjMesh& a, double omega, const double* Gbar, double* Dbar,
                           double* R1bar, hipStream_t s);
// cloth curvature + tet projection adjoints, then the per-point gather:
//   ax = D u0 + J_local(x)^T bbar;  sbar += d rhs / d spring weight;  rhs0bar += bbar;
//   colRbar += bbar (when the collision rhs entered the local step)
void launch_adj_local(const DeviceMesh& m, const AdjMesh& a, double sigLo, double sigHi, const double* x,
                      const double* bbar, const double* u0, double* aA, double* tetAdj, double* ax, double* sbar,
                      double* rhs0bar, double* colRbar, hipStream_t s);
// per-record TOI (2 where skipped), the minimum into toiBits, the first record
// attaining it into *argmin (which the caller sets to INT_MAX beforehand)
void launch_ccd_records(const DeviceMesh& m, const int4* recs, int nRecs, const double* x0, const double* x1,
                        double filter, int iters, double step, double gap, int touchPolicy, double* recToi,
                        int* argmin, hipStream_t s);
// diagnostics of the frozen choices: the ACCD exit path of the record in
// argmin[0] into argmin[5]; the active barrier records' hash added to *hash
void launch_argmin_path(const DeviceMesh& m, int nRecs, const double* x0, const double* x1, double filter, int iters,
                        double step, double gap, int touchPolicy, int* argmin, hipStream_t s);
void launch_active_hash(const DeviceMesh& m, unsigned long long* hash, hipStream_t s);
// over the active records `act` (indices into recs) of that collision pass
void launch_adj_collision(const DeviceMesh& m, const int4* recs, const int* act, int nAct, const double* x, double dhat,
                          double kappa, double dt, const double* colDbar, const double* colRbar, double* ax,
                          hipStream_t s);
void launch_adj_ccd(const DeviceMesh& m, const int4* recs, int argmin, const double* x0, const double* x1,
                    double filter, int iters, double step, double gap, int touchPolicy, const double* gscale,
                    double* ax0, double* ax1, hipStream_t s);
// elementwise helpers over the simulated rows of [3][ld] vectors
void launch_axpby(const DeviceMesh& m, double a, const double* x, double b, const double* y, double* out, hipStream_t s);
void launch_div_rows(const DeviceMesh& m, const double* x, const double* D, double* out, hipStream_t s);
void launch_mul_rows(const DeviceMesh& m, const double* x, const double* D, double* out, hipStream_t s);
void launch_dot_dir(const DeviceMesh& m, const double* a, const double* x1, const double* x0, double scale,
                    double* partials, double* out, hipStream_t s);
void launch_predict_adj(const DeviceMesh& m, double dt, const double* as, const double* rhs0bar, const double* axn,
                        double* xbar, double* vbar, hipStream_t s);
void launch_add_rows(double* dst, const double* src, int64_t n, hipStream_t s);

void launch_toi_init(const DeviceMesh& m, hipStream_t s);
void launch_ccd(const DeviceMesh& m, int nRecs, double filter, int iters, double step, double gap, int touchPolicy,
                hipStream_t s);
// CCD over m.recs on AoS positions; the minimum of workgroup b into blockMin[b]
// for b < ceil(nRecs / kBlock) (every such entry written, no atomics)
void launch_ccd_blockmin(const DeviceMesh& m, int nRecs, const double* x0Aos, const double* x1Aos, double filter,
                         int iters, double step, double gap, int touchPolicy, double* blockMin, hipStream_t s);
// writes colDiag / colRhs of every simulated point (deterministic order)
int launch_collision(DeviceMesh& m, int nRecs, double dhat, double kappa, double dt, hipStream_t s);
// k_broad.hip
// runs patch boxes, LBVH build, pair search and record expansion; returns the
// pair and record totals (host sync) and whether the capacities sufficed.
void launch_patch_boxes(const DeviceMesh& m, const double* x, const double* dx, double inflation, hipStream_t s);
void launch_lbvh(DeviceMesh& m, hipStream_t s);
// the two traversal passes: top (counts in counters[3], [6], [7]; the host
// grows the pass-1 segments when counters[3] > 0 and repeats), then the
// compaction of the work items and the subtree pass (its pair total in
// counters[0]), then the compaction of both passes' pairs into m.pairs
int traverse_top_waves(int nPatches);
int traverse_sub_cap();  // pairs per work item (at most)
void launch_traverse_top(const DeviceMesh& m, int skipStatic, hipStream_t s);
void launch_traverse_sub(const DeviceMesh& m, int skipStatic, int nItems, hipStream_t s);
void launch_pairs_compact(const DeviceMesh& m, int nItems, int nTopPairs, hipStream_t s);
void launch_rec_count(const DeviceMesh& m, int nPairs, hipStream_t s);
// adds to *count the FP64 edge-triangle crossings at m.X over the last pair
// list (both ways) and within each patch (validation)
void launch_count_crossings(const DeviceMesh& m, int nPairs, unsigned long long* count, hipStream_t s);
void launch_rec_fill(const DeviceMesh& m, int nPairs, int nPT, int nRecs, hipStream_t s);
void launch_exclusive_scan(int* data_in, int* data_out, int n, int* total_out, int* blockScratch, hipStream_t s);

}  // namespace pdipc
