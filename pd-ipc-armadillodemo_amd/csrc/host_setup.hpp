// Host-side precompute of the PD-IPC step: mesh topology, per-model constraint
// data, the PD system matrix, the R1 / R2 (M^2) Jacobi operators and the
// collision patches.  Mirrors PdIpcSimulator::initialization
// (Simulator/PD-IPC/PdIpcSimulator.cpp:34-268), setupHessMat (:270-414),
// gpuSolverInit (:563-912) and gpuCollisionInit (:1151-1504), producing flat
// arrays ready for the device.  No HIP calls: this file also builds and runs
// on a machine without a GPU.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace pdipc {

enum ModelKind { MODEL_CLOTH = 0, MODEL_TET = 1, MODEL_STATIC = 2 };

struct ModelInput {
  int kind = MODEL_CLOTH;
  std::vector<double> verts;  // 3*nv AoS
  std::vector<int> elems;     // 3*nf triangles or 4*nt tets
  double s0 = 0, s1 = 0, s2 = 0, density = 1;
  int patchDim = 10;
  std::vector<int> fixed;
  // optional per-triangle strain stiffness of a cloth: an edge's spring weight
  // is the mean over its adjacent triangles (empty: s0 on every edge)
  std::vector<double> triK;
  int nv() const { return (int)verts.size() / 3; }
};

struct StepConstants {
  double dt = 0.01, gravity = -9.8, dhat = 1e-3;
};

// Everything the device needs, in the order of the reference's device buffers.
struct HostSetup {
  int nTotal = 0, nSim = 0, nCloth = 0, nTets = 0;
  std::vector<double> X;                 // 3*nTotal AoS, initial positions
  std::vector<double> mass, invMass;     // nSim
  std::vector<double> gravY;             // nSim: gravity * mass (y component)
  std::vector<double> fixK;              // nSim: position stiffness, 0 if free
  std::vector<int> fixFlag;              // nSim: 1 if fixed
  std::vector<double> elasticDiag;       // nSim
  // cloth: per cloth vertex (reference order), ring with the vertex itself first
  std::vector<int> cCenter;              // nCloth: point id
  std::vector<int> cOff, cNbr;           // CSR over slots
  std::vector<double> cRest, cLbo, cBendIn;  // per slot
  std::vector<double> cStrain;           // per slot: spring weight of the slot's edge
  std::vector<double> cH0;               // per cloth vertex
  std::vector<int> clothOfPoint;         // nSim: cloth index or -1
  // derivative bookkeeping (used by the adjoint)
  std::vector<double> cBendOut;          // per slot: bend(i -> slot) = L_i,slot * bend * area_i
  std::vector<double> cLboIn;            // per slot s of i (k = nbr): L_k,i (the gather form of L^T)
  std::vector<int> cSlotEdge;            // per slot: cloth edge id (-1 for the vertex itself)
  std::vector<int> cSlotR1;              // per slot: R1 CSR index of (i, nbr) (-1 for the vertex itself)
  std::vector<int> edgeTriOff, edgeTriList;  // cloth edge -> cloth triangles (ids over all cloths)
  std::vector<int> clothTriBase;         // per simulated model: first cloth-triangle id (-1: not a cloth)
  int nClothTris = 0;
  // tets
  std::vector<int> tetIdx;               // 4*T
  std::vector<double> tetDminv;          // 9*T row-major Dm^{-1}
  std::vector<double> tetW;              // T: stiffness * rest volume
  std::vector<int> vtOff, vtList;        // point -> (tet*4 + corner)
  // R1 = -offdiag(sysHess), CSR without the diagonal slot
  std::vector<int> r1Off, r1Col;
  std::vector<double> r1Val;
  // M^2 pattern (CSR, ascending columns) with its R1 value and its 2-paths
  std::vector<int> m2Off, m2Col;
  std::vector<double> m2R1;              // R1(i,j) (0 if not an R1 entry)
  std::vector<int64_t> p2Off;            // per M^2 entry
  std::vector<int> p2K;                  // middle index k
  std::vector<double> p2A;               // R1(i,k) * R1(k,j)
  std::vector<int> p2Ia, p2Ib;           // R1 CSR indices of (i,k) and (j,k) of each 2-path
  std::vector<int> m2R1Idx;              // R1 CSR index of (i,j), -1 if none
  // collision
  std::vector<int> bFaces, bEdges;       // 3*F, 2*E point ids
  std::vector<int> pfOff, pfList;        // patch -> boundary faces
  std::vector<int> ppOff, ppList;        // patch -> owned points
  std::vector<int> peOff, peList;        // patch -> owned edges
  std::vector<uint8_t> patchStatic;      // 1 if the patch belongs to a static model
  std::vector<double> colMass;           // nTotal
  // fixture images (boundary.dat / patchInfo.dat) per simulated model
  std::vector<std::vector<int>> boundaryDat, patchDat;
  int64_t nnzR1 = 0, nnzR2 = 0, n2paths = 0;
};

// Runs the whole precompute.  Returns an empty string on success, otherwise an
// error message (degenerate input).
std::string build_host_setup(const std::vector<ModelInput>& sims, const std::vector<ModelInput>& statics,
                             const StepConstants& k, HostSetup& out);

constexpr int kSellLaneMajorW = 8;  // slice width of the lane-major layout

// SELL-C layout of the M^2 pattern for the Chebyshev sweep: slices of
// (64 / lanes) rows, 64 lanes x sliceW[s] entries each; lane l serves slot-row
// l / lanes, whose CSR entry c is held by lane (l / lanes) * lanes + c % lanes
// as its chunk c / lanes.  A slice is chunk-major, entry `chunk` of lane l at
//   sliceOff[s] + chunk * 64 + l
// (a wave's loads of one chunk are one run of 64 entries), except when every
// slice has width kSellLaneMajorW: then slices are lane-major, at
//   sliceOff[s] + l * kSellLaneMajorW + chunk
// (a lane's entries are one aligned run, read with 16-byte loads).
// Padding entries point at their own row with no 2-paths.
struct SellLayout {
  int lanes = 4, rowsPerSlice = 16, nSlices = 0;
  int uniformW = 0;               // > 0: every slice padded to this width
  // non-uniform layouts: slot -> row (rows sorted by width within windows of
  // kSellSortSlices slices; slots past the last row map to themselves) and its
  // inverse over the rows; both empty for the identity order
  std::vector<int> perm, slotOf;
  std::vector<int64_t> sliceOff;   // nSlices + 1
  std::vector<int> sliceW;         // nSlices: entries per lane
  bool laneMajor() const { return uniformW == kSellLaneMajorW; }
  // the SELL entry `chunk` of `lane` in slice s, and the lane holding entry e
  int64_t entry(int s, int lane, int chunk) const {
    return sliceOff[s] + (laneMajor() ? (int64_t)lane * sliceW[s] + chunk : (int64_t)chunk * 64 + lane);
  }
  int lane_of(int s, int64_t e) const {
    const int64_t d = e - sliceOff[s];
    return (int)(laneMajor() ? d / sliceW[s] : d % 64);
  }
  std::vector<int> col;            // per SELL entry
  // per SELL entry: col - the entry's slot index (its row without perm; padding
  // rows included);
  // filled only when every offset fits in 16 bits (col16ok)
  std::vector<int16_t> col16;
  bool col16ok = false;
  std::vector<int> rowFlag;        // row | 0x80000000 on the diagonal entry
  std::vector<double> r1;          // per SELL entry
  std::vector<int64_t> pOff;       // per SELL entry + 1
  std::vector<int> pK;
  std::vector<double> pA;
  std::vector<int> r1Idx;          // per SELL entry: R1 index of (i,j), -1 if none or padding
  std::vector<int> pIa, pIb;       // per 2-path: R1 indices of its two factors
  // LDS staging of the sweeps' gathered operand, per tile of kTileSlices
  // consecutive slices (one workgroup): the sorted union of the columns of the
  // tile's entries (padding included) as 16-bit offsets from the tile's
  // smallest column, and per entry its index in that union.  stageOk: every
  // tile's union spans at most 65536 columns and holds at most kStageMaxCols.
  bool stageOk = false;
  int stageMax = 0;                // largest union
  std::vector<int> tileOff;        // nTiles + 1: start of each tile's union in tileCol
  std::vector<int> tileBase;       // nTiles: smallest column of the tile
  std::vector<uint16_t> tileCol;   // union columns - tileBase
  std::vector<uint16_t> loc;       // per SELL entry: index in its tile's union
  // Narrow staging (lane-major layouts with 8 lanes per row; the FP32 and
  // FP64 sweeps): tiles of kNarrowTileSlices slices whose
  // unions hold at most kNarrowMaxCols columns, so an entry's union index is
  // one byte.  narrowOk: every narrow tile fits.
  bool narrowOk = false;
  int narrowMax = 0;
  std::vector<int> ntileOff, ntileBase;
  std::vector<uint16_t> ntileCol;
  std::vector<uint8_t> loc8;
};
constexpr int kSellSortSlices = 2;  // slices per row-sorting window
constexpr int kTileSlices = 4;      // slices per staged tile (a 256-lane workgroup)
constexpr int kStageMaxCols = 1024; // largest staged union (16 KiB of FP32 rows, 32 KiB of FP64 rows)
constexpr int kNarrowTileSlices = 2;  // slices per narrow tile (a 128-lane workgroup)
constexpr int kNarrowMaxCols = 256;   // largest narrow union (8-bit indices)
void build_sell(const HostSetup& hs, int lanes, SellLayout& out);

}  // namespace pdipc
