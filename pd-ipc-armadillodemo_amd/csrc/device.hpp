// Internal: device-side data of one simulator handle, and the launch wrappers
// each kernel file provides.  Positions and every 3-vector field are stored SoA
// with leading dimension `ld` (component c of point i at [c * ld + i]); the
// host boundary converts from / to the reference's AoS xyz layout.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pdipc {

constexpr int kWave = 64;

struct DeviceMesh {
  int nTotal = 0, nSim = 0, nCloth = 0, nTets = 0;
  int ld = 0;  // leading dimension of SoA vectors (>= nTotal, multiple of 64)
  // state and work vectors, SoA [3][ld] doubles
  double *X = nullptr, *Xold = nullptr, *X0 = nullptr, *Dir = nullptr, *Pre = nullptr, *Next = nullptr, *V = nullptr;
  double *initRhs = nullptr, *sysRhs = nullptr, *firstTerm = nullptr, *colRhs = nullptr, *clothA = nullptr, *Xrest = nullptr;
  // FP32-split work vectors, SoA [3][ld] floats
  float *dPre = nullptr, *dCur = nullptr, *dNext = nullptr, *r0 = nullptr;
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
  int64_t sellEntries = 0;
  int64_t* sliceOff = nullptr;
  int* sliceW = nullptr;
  int *m2Col = nullptr, *m2RowFlag = nullptr;
  double *m2R1 = nullptr, *m2Val = nullptr;
  float* m2ValF = nullptr;
  int64_t* p2Off = nullptr;
  int* p2K = nullptr;
  double* p2A = nullptr;
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
  double* nodeBox = nullptr;       // SoA [6][2P-1]
  int2* pairs = nullptr;
  int pairCap = 0;
  int *recCount = nullptr, *recOff = nullptr;
  int4* recs = nullptr;            // EE records store ~a in .x
  int recCap = 0;
  int* scanBlock = nullptr;        // scan scratch
  int scanBlockCap = 0;
  // scalars
  unsigned long long* toiBits = nullptr;
  double* partials = nullptr;      // reduction partials
  double* residual = nullptr;
  int* counters = nullptr;         // [0] = pairs, [1] = records, [2] = active pairs
  int nPartials = 0;
};

// ---- kernel launch wrappers (all asynchronous on `s`) -------------------------
// k_step.hip
void launch_predict(const DeviceMesh& m, double dt, hipStream_t s);
void launch_diag(const DeviceMesh& m, int withCollision, hipStream_t s);
void launch_direction(const DeviceMesh& m, hipStream_t s);
void launch_clamp_residual(const DeviceMesh& m, double toiScale, hipStream_t s);
void launch_velocity(const DeviceMesh& m, double dt, hipStream_t s);
void launch_zero(double* p, int64_t n, hipStream_t s);
void launch_copy(double* dst, const double* src, int64_t n, hipStream_t s);
// k_local.hip
void launch_local(const DeviceMesh& m, int precision, double sigLo, double sigHi, int withCollision, hipStream_t s);
// k_solve.hip
void launch_operators(const DeviceMesh& m, double omega, int precision, hipStream_t s);
void launch_solve(const DeviceMesh& m, double omega, double rho, int chebIters, int precision, hipStream_t s);
// one Chebyshev sweep of the given precision (measurement support)
void launch_one_sweep(const DeviceMesh& m, int precision, hipStream_t s);
// k_contact.hip
void launch_toi_init(const DeviceMesh& m, hipStream_t s);
void launch_ccd(const DeviceMesh& m, int nRecs, double filter, int iters, double step, double gap, int touchPolicy,
                hipStream_t s);
void launch_collision(const DeviceMesh& m, int nRecs, double dhat, double kappa, double dt, hipStream_t s);
// k_broad.hip
// runs patch boxes, LBVH build, pair search and record expansion; returns the
// pair and record totals (host sync) and whether the capacities sufficed.
void launch_patch_boxes(const DeviceMesh& m, const double* x, const double* dx, double inflation, hipStream_t s);
void launch_lbvh(DeviceMesh& m, hipStream_t s);
// appends overlapping patch pairs, count in counters[0]
void launch_pairs(const DeviceMesh& m, int skipStatic, hipStream_t s);
void launch_rec_count(const DeviceMesh& m, int nPairs, hipStream_t s);
void launch_rec_fill(const DeviceMesh& m, int nPairs, int nRecs, hipStream_t s);
void launch_exclusive_scan(int* data_in, int* data_out, int n, int* total_out, int* blockScratch, hipStream_t s);

}  // namespace pdipc
