// TEST INFRASTRUCTURE ONLY -- the oracle.
//
// CPU restatement of the reference PD-IPC time step (lanlei/PD-IPC-ArmadilloDemo,
// Simulator/PD-IPC/*).  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library; the product (the HIP library in
// pd-ipc-armadillodemo_amd/) never links, loads or calls it.
//
// It is written to read like the reference: per-vertex / per-element loops over
// AoS xyz arrays, vectors of vectors for the flattened 2-D tables, and the same
// order of operations as each reference kernel, so that a reader can check it
// line by line against the cited source.  The scalar type is a template
// parameter: Real = double is the FP64 oracle, Real = float mimics the
// reference's own FP32 build (DataCore.h:7 leaves USE_DOUBLE_PRECISION off).
//
// Parity status of each function is stated in its comment:
//   "pinned"   = checked by tests against a reference fixture, a closed form,
//                an invariant or brute force;
//   "unpinned" = no independent check exists (stated where it applies).
//
// Deliberate deviations from the reference (the GPU library makes the same ones):
//   D1 inflation radius: the reference's per-step broad phase reads an
//      uninitialised member (PdIpcSimulator.cpp:1154 shadows PdIpcSimulator.h:302);
//      we use 0.5*dHat in both places, the evident intent.
//   D2 static points have zero displacement in the broad phase (the reference
//      reads never-written device memory there, PdIpcSimulator.cpp:1715-1723).
//   D3 degenerate primitives (triangle area or edge cross product below 1e-24)
//      fall back to the minimum over all edges, where the reference leaves the
//      barycentric outputs uninitialised (PdIpcGpuFunc.cu:2194-2230,2255-2286).
//   D4 touching pairs in CCD (distance exactly 0) give TOI 0 when
//      touch_policy = 1; the reference produces NaN and ignores the pair
//      (PdIpcGpuFunc.cu:2499,2669).  touch_policy = 0 keeps the reference rule.
//   D5 the tet boundary (volumetric models) is listed in sorted-key order where
//      the reference iterates an unordered_map (PdIpcVolumetricModel.h:53).
//   D6 the 3x3 SVD is a cyclic Jacobi eigen-solve of F^T F run to convergence
//      followed by Gram-Schmidt (the structure of McAdams et al. 2011, whose
//      4-sweep approximate-Givens variant the reference uses, CudaSVD.cu:291-316).

#include <omp.h>

// Planted mistakes for the mutation tests (tests/test_oracle_mutants.py): a
// build with -DORC_MUTANT=k carries misreading k of the reference, and the
// closed-form tests must fail on it.  The default build has none.
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif
#define ORC_MUT(k) (ORC_MUTANT == (k))

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <limits>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <type_traits>
#include <vector>

namespace orc {

using std::acos;
using std::fabs;
using std::log;
using std::sqrt;
using std::tan;

// ---------------------------------------------------------------------------
// Forward-mode dual numbers: a value and K tangents.  Sim<Dual<K>> runs the
// same step as Sim<double> -- every value is computed by the same operations in
// the same order, so the value part is bit-identical -- and carries K
// directional derivatives of every quantity along.  This is the CPU reference
// for the derivative of the step: the chain rule is applied mechanically to
// the restatement, independently of the hand-written adjoint it checks.
// ---------------------------------------------------------------------------
template <int K>
struct Dual {
  double v;
  double d[K];
  Dual() : v(0) { for (int k = 0; k < K; k++) d[k] = 0; }
  Dual(double a) : v(a) { for (int k = 0; k < K; k++) d[k] = 0; }
  explicit operator double() const { return v; }
  explicit operator float() const { return (float)v; }
  Dual& operator+=(const Dual& o) { v += o.v; for (int k = 0; k < K; k++) d[k] += o.d[k]; return *this; }
  Dual& operator-=(const Dual& o) { v -= o.v; for (int k = 0; k < K; k++) d[k] -= o.d[k]; return *this; }
  Dual& operator*=(const Dual& o) { *this = *this * o; return *this; }
  Dual& operator/=(const Dual& o) { *this = *this / o; return *this; }
  friend Dual operator+(Dual a, const Dual& b) { return a += b; }
  friend Dual operator-(Dual a, const Dual& b) { return a -= b; }
  friend Dual operator-(const Dual& a) { Dual r; r.v = -a.v; for (int k = 0; k < K; k++) r.d[k] = -a.d[k]; return r; }
  friend Dual operator*(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v * b.v;
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
  }
  friend Dual operator/(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v / b.v;
    for (int k = 0; k < K; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
    return r;
  }
  friend bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
  friend bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
  friend bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
  friend bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
  friend bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
  friend bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }
  friend Dual sqrt(const Dual& a) {
    Dual r;
    r.v = std::sqrt(a.v);
    for (int k = 0; k < K; k++) r.d[k] = r.v > 0 ? a.d[k] / (2 * r.v) : 0.0;
    return r;
  }
  friend Dual fabs(const Dual& a) { return a.v < 0 ? -a : a; }
  friend Dual log(const Dual& a) {
    Dual r;
    r.v = std::log(a.v);
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] / a.v;
    return r;
  }
  friend Dual acos(const Dual& a) {
    Dual r;
    r.v = std::acos(a.v);
    double g = -1.0 / std::sqrt(1.0 - a.v * a.v);
    for (int k = 0; k < K; k++) r.d[k] = g * a.d[k];
    return r;
  }
  friend Dual tan(const Dual& a) {
    Dual r;
    r.v = std::tan(a.v);
    for (int k = 0; k < K; k++) r.d[k] = (1.0 + r.v * r.v) * a.d[k];
    return r;
  }
};
constexpr int kDualDirs = 8;
using DualD = Dual<kDualDirs>;

}  // namespace orc

namespace std {
template <int K>
struct numeric_limits<orc::Dual<K>> {
  static orc::Dual<K> max() { return orc::Dual<K>(numeric_limits<double>::max()); }
};
}  // namespace std

namespace orc {

// ---------------------------------------------------------------------------
// parameters (flat double array; order shared with oracle/oracle.py)
// ---------------------------------------------------------------------------
enum ParamIndex {
  P_DT = 0, P_GRAVITY, P_DHAT, P_KAPPA, P_LG_ITERS, P_CHEB_ITERS, P_MAX_OUTER,
  P_RESIDUAL_TOL, P_OMEGA, P_RHO, P_SIG_LO, P_SIG_HI, P_CCD_FILTER, P_TOI_SCALE,
  P_CCD_ITERS, P_CCD_STEP, P_CCD_GAP, P_STATIC_MASS, P_INFLATION,
  P_SKIP_STATIC_PAIRS, P_TOUCH_POLICY, P_COUNT
};

struct Params {
  double v[P_COUNT];
  Params() {
    v[P_DT] = 0.01; v[P_GRAVITY] = -9.8; v[P_DHAT] = 1e-3; v[P_KAPPA] = 1e4;
    v[P_LG_ITERS] = 10; v[P_CHEB_ITERS] = 15; v[P_MAX_OUTER] = 30;
    v[P_RESIDUAL_TOL] = 1e-5; v[P_OMEGA] = 2.0 / 3.0; v[P_RHO] = 0.9992;
    v[P_SIG_LO] = 0.999; v[P_SIG_HI] = 1.001; v[P_CCD_FILTER] = 5e-3;
    v[P_TOI_SCALE] = 0.98; v[P_CCD_ITERS] = 10; v[P_CCD_STEP] = 0.9;
    v[P_CCD_GAP] = 0.1; v[P_STATIC_MASS] = 1e4; v[P_INFLATION] = -1.0;
    v[P_SKIP_STATIC_PAIRS] = 0; v[P_TOUCH_POLICY] = 1;
  }
};

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
template <class R> static inline R min2(R a, R b) { return a < b ? a : b; }
template <class R> static inline R max2(R a, R b) { return a > b ? a : b; }
template <class R> static inline R min3(R a, R b, R c) { return a < b ? (a < c ? a : c) : (b < c ? b : c); }
template <class R> static inline R max3(R a, R b, R c) { return a > b ? (a > c ? a : c) : (b > c ? b : c); }
template <class R> static inline void sub3(const R* a, const R* b, R* r) { r[0] = a[0] - b[0]; r[1] = a[1] - b[1]; r[2] = a[2] - b[2]; }
template <class R> static inline R dot3(const R* a, const R* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
template <class R> static inline void cross3(const R* a, const R* b, R* r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

// ---------------------------------------------------------------------------
// surface-mesh topology: BaseSurfaceMesh::readMeshFromObjFormat edge build
// (Model/TriangleMesh/BaseSurfaceMesh.cpp:219-277).  Pinned by boundary.dat.
// ---------------------------------------------------------------------------
struct SurfaceTopo {
  std::vector<std::array<int, 2>> edges;          // sorted (v2IndComp), unique
  std::vector<std::vector<int>> pointEdges;       // pointSharedEdges
  std::vector<std::vector<int>> edgeFaces;        // edgeSharedFaces
};

static SurfaceTopo build_surface_topo(int nv, const std::vector<int>& F) {
  SurfaceTopo t;
  int nf = (int)F.size() / 3;
  std::vector<std::array<int, 2>> list;
  list.reserve(3 * nf);
  for (int j = 0; j < nf; j++)
    for (int k = 0; k < 3; k++) {
      int a = F[3 * j + k % 3], b = F[3 * j + (k + 1) % 3];
      list.push_back(a < b ? std::array<int, 2>{a, b} : std::array<int, 2>{b, a});
    }
  std::sort(list.begin(), list.end());
  list.erase(std::unique(list.begin(), list.end()), list.end());
  t.edges = list;
  t.pointEdges.assign(nv, {});
  t.edgeFaces.assign(list.size(), {});
  std::map<std::array<int, 2>, int> ids;
  for (int i = 0; i < (int)list.size(); i++) {
    t.pointEdges[list[i][0]].push_back(i);
    t.pointEdges[list[i][1]].push_back(i);
    ids[list[i]] = i;
  }
  for (int i = 0; i < nf; i++)
    for (int j = 0; j < 3; j++) {
      int a = F[3 * i + j % 3], b = F[3 * i + (j + 1) % 3];
      std::array<int, 2> e = a < b ? std::array<int, 2>{a, b} : std::array<int, 2>{b, a};
      t.edgeFaces[ids[e]].push_back(i);
    }
  return t;
}

// ---------------------------------------------------------------------------
// model input + model-local collision topology
// ---------------------------------------------------------------------------
enum Kind { KIND_CLOTH = 0, KIND_TET = 1, KIND_STATIC = 2 };

struct Model {
  int kind = 0;
  int nv = 0;
  std::vector<double> V;   // 3*nv
  std::vector<int> F;      // surface triangles (cloth / static)
  std::vector<int> T;      // tets
  double s0 = 0, s1 = 0, s2 = 0, density = 1;
  int patchDim = 10;
  // optional per-triangle strain stiffness (cloth): an edge's spring weight is
  // the mean over its adjacent triangles (a documented extension; the
  // reference uses one weight per model)
  std::vector<double> triK;
  std::vector<int> fixedIds;
  // model-local boundary (PdIpcModel.h:22-37)
  std::vector<int> bPoints;
  std::vector<std::array<int, 2>> bEdges;
  std::vector<std::array<int, 3>> bFaces;
  std::vector<std::vector<int>> patches;              // collisionPatches
  std::vector<std::vector<int>> facesSharedEdges;     // boundaryFacesSharedEdges
  std::vector<std::vector<int>> facesSharedPoints;    // boundaryFacesSharedPoints (point ids)
  SurfaceTopo topo;                                   // cloth / static only
};

// PdIpcClothModel::registerBoundaryInfo (PdIpcClothModel.h:24-43).  Pinned by boundary.dat.
static void cloth_register_boundary(Model& m) {
  m.bPoints.resize(m.nv);
  for (int i = 0; i < m.nv; i++) m.bPoints[i] = i;
  m.bEdges = m.topo.edges;
  int nf = (int)m.F.size() / 3;
  m.bFaces.resize(nf);
  for (int i = 0; i < nf; i++) m.bFaces[i] = {m.F[3 * i], m.F[3 * i + 1], m.F[3 * i + 2]};
}

// PdIpcVolumetricModel::registerBoundaryInfo (PdIpcVolumetricModel.h:27-113),
// with deviation D5 (sorted-key order).  Unpinned beyond the invariants tested
// (every boundary face appears once, outward orientation).
static void tet_register_boundary(Model& m) {
  static const int eb[4][3] = {{3, 1, 0}, {3, 2, 1}, {3, 0, 2}, {2, 1, 0}};
  std::map<std::array<int, 3>, std::vector<std::array<int, 4>>> faceMap;
  int nt = (int)m.T.size() / 4;
  for (int e = 0; e < nt; e++)
    for (int k = 0; k < 4; k++) {
      std::array<int, 3> f = {m.T[4 * e + eb[k][0]], m.T[4 * e + eb[k][1]], m.T[4 * e + eb[k][2]]};
      std::array<int, 3> s = f;
      std::sort(s.begin(), s.end());
      faceMap[s].push_back({f[0], f[1], f[2], e});
    }
  for (auto& it : faceMap) {
    if (it.second.size() != 1) continue;
    auto d = it.second[0];
    std::array<int, 3> face = {d[0], d[1], d[2]};
    const int* ti = &m.T[4 * d[3]];
    double cent[3] = {0, 0, 0}, fc[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++) {
      for (int j = 0; j < 4; j++) cent[k] += m.V[3 * ti[j] + k];
      cent[k] /= 4;
      for (int j = 0; j < 3; j++) fc[k] += m.V[3 * face[j] + k];
      fc[k] /= 3;
    }
    double e1[3], e2[3], n[3], cd[3];
    sub3(&m.V[3 * face[1]], &m.V[3 * face[0]], e1);
    sub3(&m.V[3 * face[2]], &m.V[3 * face[0]], e2);
    cross3(e1, e2, n);
    sub3(cent, fc, cd);
    if (dot3(cd, n) > 0) face = {d[0], d[2], d[1]};
    m.bFaces.push_back(face);
  }
  std::set<int> ps;
  std::set<std::array<int, 2>> es;
  for (auto& f : m.bFaces)
    for (int k = 0; k < 3; k++) {
      ps.insert(f[k]);
      int a = f[k], b = f[(k + 1) % 3];
      es.insert(a < b ? std::array<int, 2>{a, b} : std::array<int, 2>{b, a});
    }
  m.bPoints.assign(ps.begin(), ps.end());
  m.bEdges.assign(es.begin(), es.end());
}

// PdIpcClothModel::setCollisionPatches (PdIpcClothModel.h:45-140).  The
// reference computes in float (qeal = float); so do we, always.  Pinned by
// patchInfo.dat.  When no extent is strictly smallest the reference leaves
// (u, v) uninitialised; we then keep the last assignment the chain would make
// with u = 0, v = 2 (the vertical axis y excluded), a stated deviation.
static void cloth_set_patches(Model& m) {
  float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  auto P = [&](int i, int k) { return (float)m.V[3 * i + k]; };
  for (int i : m.bPoints)
    for (int k = 0; k < 3; k++) {
      mn[k] = min2(mn[k], P(i, k));
      mx[k] = max2(mx[k], P(i, k));
    }
  float sq = sqrt((float)m.nv);
  int len = sq < (float)m.patchDim ? (int)sq : m.patchDim;
  float size[3];
  for (int k = 0; k < 3; k++) size[k] = mx[k] - mn[k];
  int u = 0, v = 2;
  if (size[0] < size[1] && size[0] < size[2]) { u = 1; v = 2; }
  if (size[1] < size[0] && size[1] < size[2]) { u = 0; v = 2; }
  if (size[2] < size[0] && size[2] < size[1]) { u = 0; v = 1; }
  int gsize[3];
  for (int k = 0; k < 3; k++) gsize[k] = (k == u || k == v) ? len : 1;
  struct B2 { float lo[2], hi[2]; };
  std::vector<B2> boxes;
  float du = size[u] / len, dv = size[v] / len;
  for (int r = 0; r < gsize[u]; r++)
    for (int c = 0; c < gsize[v]; c++) {
      B2 b;
      b.lo[0] = mn[u] + r * du;
      b.lo[1] = mn[v] + c * dv;
      b.hi[0] = b.lo[0] + du;
      b.hi[1] = b.lo[1] + dv;
      boxes.push_back(b);
    }
  std::vector<std::vector<int>> pocket(boxes.size());
  for (int i = 0; i < (int)m.bFaces.size(); i++) {
    float cent[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++)
      for (int j = 0; j < 3; j++) cent[j] += P(m.bFaces[i][k], j);
    for (int j = 0; j < 3; j++) cent[j] /= 3;
    for (int j = 0; j < (int)boxes.size(); j++)
      if (cent[u] >= boxes[j].lo[0] && cent[u] <= boxes[j].hi[0] && cent[v] >= boxes[j].lo[1] &&
          cent[v] <= boxes[j].hi[1]) {
        pocket[j].push_back(i);
        break;
      }
  }
  for (auto& p : pocket)
    if (!p.empty()) m.patches.push_back(p);
}

// PdIpcVolumetricModel::setCollisionPatches (PdIpcVolumetricModel.h:115-181), float.
static void tet_set_patches(Model& m) {
  float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  auto P = [&](int i, int k) { return (float)m.V[3 * i + k]; };
  for (int i : m.bPoints)
    for (int k = 0; k < 3; k++) {
      mn[k] = min2(mn[k], P(i, k));
      mx[k] = max2(mx[k], P(i, k));
    }
  int len = m.patchDim;
  float size[3];
  for (int k = 0; k < 3; k++) size[k] = mx[k] - mn[k];
  struct B3 { float lo[3], hi[3]; };
  std::vector<B3> boxes;
  for (int s = 0; s < len; s++)
    for (int r = 0; r < len; r++)
      for (int c = 0; c < len; c++) {
        B3 b;
        int idx[3] = {s, r, c};
        for (int k = 0; k < 3; k++) {
          b.lo[k] = mn[k] + idx[k] * (size[k] / len);
          b.hi[k] = b.lo[k] + size[k] / len;
        }
        boxes.push_back(b);
      }
  std::vector<std::vector<int>> pocket(boxes.size());
  for (int i = 0; i < (int)m.bFaces.size(); i++) {
    float cent[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++)
      for (int j = 0; j < 3; j++) cent[j] += P(m.bFaces[i][k], j);
    for (int j = 0; j < 3; j++) cent[j] /= 3;
    for (int j = 0; j < (int)boxes.size(); j++) {
      const B3& b = boxes[j];
      if (cent[0] >= b.lo[0] && cent[0] <= b.hi[0] && cent[1] >= b.lo[1] && cent[1] <= b.hi[1] &&
          cent[2] >= b.lo[2] && cent[2] <= b.hi[2]) {
        pocket[j].push_back(i);
        break;
      }
    }
  }
  for (auto& p : pocket)
    if (!p.empty()) m.patches.push_back(p);
}

// PdIpcModel::postProcessBoundaryInfo (PdIpcModel.h:84-172): every boundary
// edge and point is owned by exactly one incident face -- the one with the
// fewest owned items so far, first in list order on ties.  Pinned indirectly
// through the candidate sets (brute-force comparison) and the fixture patches.
static void post_process_boundary(Model& m) {
  int np = (int)m.bPoints.size(), ne = (int)m.bEdges.size(), nf = (int)m.bFaces.size();
  std::map<int, int> pid;
  for (int i = 0; i < np; i++) pid[m.bPoints[i]] = i;
  std::vector<std::vector<int>> pointFaces(np);
  for (int i = 0; i < nf; i++)
    for (int k = 0; k < 3; k++) pointFaces[pid[m.bFaces[i][k]]].push_back(i);
  std::map<std::array<int, 2>, int> eid;
  for (int i = 0; i < ne; i++) eid[m.bEdges[i]] = i;
  std::vector<std::vector<int>> edgeFaces(ne);
  for (int i = 0; i < nf; i++) {
    const auto& f = m.bFaces[i];
    int pairs[3][2] = {{f[0], f[1]}, {f[1], f[2]}, {f[2], f[0]}};
    for (auto& p : pairs) {
      std::array<int, 2> e = p[0] < p[1] ? std::array<int, 2>{p[0], p[1]} : std::array<int, 2>{p[1], p[0]};
      edgeFaces[eid[e]].push_back(i);
    }
  }
  m.facesSharedEdges.assign(nf, {});
  for (int i = 0; i < ne; i++) {
    int best = edgeFaces[i][0];
    size_t bs = m.facesSharedEdges[best].size();
    for (size_t k = 1; k < edgeFaces[i].size(); k++) {
      int f = edgeFaces[i][k];
      if (m.facesSharedEdges[f].size() < bs) { best = f; bs = m.facesSharedEdges[f].size(); }
    }
    m.facesSharedEdges[best].push_back(i);
  }
  m.facesSharedPoints.assign(nf, {});
  for (int i = 0; i < np; i++) {
    int best = pointFaces[i][0];
    size_t bs = m.facesSharedPoints[best].size();
    for (size_t k = 1; k < pointFaces[i].size(); k++) {
      int f = pointFaces[i][k];
      if (m.facesSharedPoints[f].size() < bs) { best = f; bs = m.facesSharedPoints[f].size(); }
    }
    m.facesSharedPoints[best].push_back(m.bPoints[i]);
  }
}

// ---------------------------------------------------------------------------
// 3x3 SVD (deviation D6).  Returns rotations U, V and sig[0] >= sig[1] >= |sig[2]|
// with the sign of det F carried by sig[2], F = U diag(sig) V^T, all row-major.
// Pinned against LAPACK (numpy) projections in tests/test_oracle_math.py.
// ---------------------------------------------------------------------------
template <class R>
static void svd3(const R* F, R* U, R* sig, R* V) {
  R A[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) A[i][j] = F[0 * 3 + i] * F[0 * 3 + j] + F[1 * 3 + i] * F[1 * 3 + j] + F[2 * 3 + i] * F[2 * 3 + j];
  R W[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  const int pq[3][2] = {{0, 1}, {1, 2}, {0, 2}};
  for (int sweep = 0; sweep < 30; sweep++) {
    R off = A[0][1] * A[0][1] + A[1][2] * A[1][2] + A[0][2] * A[0][2];
    R dia = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
    if (!(off > dia * (R)1e-34) || off == 0) break;
    for (auto& e : pq) {
      int p = e[0], q = e[1];
      R apq = A[p][q];
      if (apq == 0) continue;
      R tau = (A[q][q] - A[p][p]) / (2 * apq);
      R t = (tau >= 0 ? (R)1 : (R)-1) / (fabs(tau) + sqrt((R)1 + tau * tau));
      R c = (R)1 / sqrt((R)1 + t * t), s = t * c;
      for (int k = 0; k < 3; k++) {  // A <- A J  (columns p, q)
        R akp = A[k][p], akq = A[k][q];
        A[k][p] = c * akp - s * akq;
        A[k][q] = s * akp + c * akq;
      }
      for (int k = 0; k < 3; k++) {  // A <- J^T A  (rows p, q)
        R apk = A[p][k], aqk = A[q][k];
        A[p][k] = c * apk - s * aqk;
        A[q][k] = s * apk + c * aqk;
      }
      for (int k = 0; k < 3; k++) {  // W <- W J
        R wkp = W[k][p], wkq = W[k][q];
        W[k][p] = c * wkp - s * wkq;
        W[k][q] = s * wkp + c * wkq;
      }
    }
  }
  R lam[3] = {A[0][0], A[1][1], A[2][2]};
  int ord[3] = {0, 1, 2};
  // sort descending (stable on ties)
  for (int i = 0; i < 3; i++)
    for (int j = i + 1; j < 3; j++)
      if (lam[ord[j]] > lam[ord[i]]) std::swap(ord[i], ord[j]);
  R Vm[3][3];
  for (int k = 0; k < 3; k++)
    for (int c = 0; c < 3; c++) Vm[k][c] = W[k][ord[c]];
  R detV = Vm[0][0] * (Vm[1][1] * Vm[2][2] - Vm[1][2] * Vm[2][1]) - Vm[0][1] * (Vm[1][0] * Vm[2][2] - Vm[1][2] * Vm[2][0]) +
           Vm[0][2] * (Vm[1][0] * Vm[2][1] - Vm[1][1] * Vm[2][0]);
  if (detV < 0)
    for (int k = 0; k < 3; k++) Vm[k][2] = -Vm[k][2];
  R b[3][3];  // columns of F V
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) b[c][r] = F[r * 3 + 0] * Vm[0][c] + F[r * 3 + 1] * Vm[1][c] + F[r * 3 + 2] * Vm[2][c];
  R u0[3], u1[3], u2[3];
  R s0 = sqrt(dot3(b[0], b[0]));
  if (s0 > (R)1e-30) {
    for (int k = 0; k < 3; k++) u0[k] = b[0][k] / s0;
  } else {
    u0[0] = 1; u0[1] = 0; u0[2] = 0;
  }
  R pr = dot3(u0, b[1]);
  for (int k = 0; k < 3; k++) u1[k] = b[1][k] - pr * u0[k];
  R s1 = sqrt(dot3(u1, u1));
  if (s1 > (R)1e-30 * (s0 > 1 ? s0 : (R)1)) {
    for (int k = 0; k < 3; k++) u1[k] /= s1;
  } else {
    // any unit vector orthogonal to u0
    R t[3] = {0, 0, 0};
    int mi = fabs(u0[0]) < fabs(u0[1]) ? (fabs(u0[0]) < fabs(u0[2]) ? 0 : 2) : (fabs(u0[1]) < fabs(u0[2]) ? 1 : 2);
    t[mi] = 1;
    cross3(u0, t, u1);
    R n = sqrt(dot3(u1, u1));
    for (int k = 0; k < 3; k++) u1[k] /= n;
    s1 = 0;
  }
  cross3(u0, u1, u2);
  R s2 = dot3(u2, b[2]);
  for (int r = 0; r < 3; r++) {
    U[r * 3 + 0] = u0[r]; U[r * 3 + 1] = u1[r]; U[r * 3 + 2] = u2[r];
    for (int c = 0; c < 3; c++) V[r * 3 + c] = Vm[r][c];
  }
  sig[0] = s0; sig[1] = s1; sig[2] = s2;
}

// Strain projection of solveVolumetricStrainConstraintKernel
// (PdIpcGpuFunc.cu:490-551): det<0 flips the last singular value and re-sorts
// the values (not the vectors), then clamp (PdIpcGpuFunc.cu:307-316), then
// P = U diag(sigma') V^T.  F and P row-major.
template <class R>
static void strain_projection(const R* F, R lo, R hi, R* P) {
  R U[9], s[3], V[9];
  svd3<R>(F, U, s, V);
  R det = s[0] * s[1] * s[2];
  if (det < 0) {
    s[2] *= -1;
    R high = s[0], mid = s[1], low = s[2];
    if (mid < low) std::swap(mid, low);
    if (high < low) std::swap(high, low);
    if (high < mid) std::swap(mid, high);
    s[0] = high; s[1] = mid; s[2] = low;
  }
  R sn[3];
  for (int k = 0; k < 3; k++) {
    R r = s[k] < hi ? s[k] : hi;
    if (r < lo) r = lo;
    sn[k] = r;
  }
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) P[r * 3 + c] = U[r * 3 + 0] * sn[0] * V[c * 3 + 0] + U[r * 3 + 1] * sn[1] * V[c * 3 + 1] + U[r * 3 + 2] * sn[2] * V[c * 3 + 2];
}

// ---------------------------------------------------------------------------
// primitive distances (GeometryDistance, PdIpcGpuFunc.cu:2152-2292) with D3.
// Pinned by closed forms and dense-sampling brute force (tests).
// ---------------------------------------------------------------------------
template <class R>
static R vertex_edge_sqd(const R* v, const R* ea, const R* eb, R* r, R* dir) {
  R va[3], ba[3];
  sub3(v, ea, va);
  sub3(eb, ea, ba);
  R va_ba = dot3(va, ba), ba_ba = dot3(ba, ba);
  if (va_ba < 0) *r = 0;
  else if (va_ba > ba_ba) *r = 1;
  else *r = va_ba / ba_ba;
  for (int k = 0; k < 3; k++) dir[k] = v[k] - (ea[k] * ((R)1 - *r) + eb[k] * *r);
  return dot3(dir, dir);
}

template <class R>
static R edge_edge_sqd(const R* xi, const R* xj, const R* xa, const R* xb, R* r, R* s, R* dir) {
  R xba[3], xji[3], xai[3], tmp[3];
  sub3(xb, xa, xba);
  sub3(xj, xi, xji);
  sub3(xa, xi, xai);
  cross3(xji, xba, dir);
  R nn = dot3(dir, dir);
  cross3(xai, xji, tmp);
  R w_aiji = dot3(dir, tmp);
  cross3(xai, xba, tmp);
  R w_aiba = dot3(dir, tmp);
  if (nn > (R)1e-24 && w_aiji >= 0 && w_aiji <= nn && w_aiba >= 0 && w_aiba <= nn) {
    *r = w_aiba / nn;
    *s = w_aiji / nn;
  } else {
    bool degen = !(nn > (R)1e-24);
    R best = (R)999999999, d, v, N[3];
    *r = 0; *s = 0;
    if ((degen || w_aiba < 0) && (d = vertex_edge_sqd(xi, xa, xb, &v, N)) < best) { best = d; *r = 0; *s = v; }
    if ((degen || w_aiba > nn) && (d = vertex_edge_sqd(xj, xa, xb, &v, N)) < best) { best = d; *r = 1; *s = v; }
    if ((degen || w_aiji < 0) && (d = vertex_edge_sqd(xa, xi, xj, &v, N)) < best) { best = d; *r = v; *s = 0; }
    if ((degen || w_aiji > nn) && (d = vertex_edge_sqd(xb, xi, xj, &v, N)) < best) { best = d; *r = v; *s = 1; }
  }
  for (int k = 0; k < 3; k++) dir[k] = xi[k] * ((R)1 - *r) + xj[k] * *r - xa[k] * ((R)1 - *s) - xb[k] * *s;
  return dot3(dir, dir);
}

template <class R>
static R vertex_triangle_sqd(const R* xi, const R* xa, const R* xb, const R* xc, R* ba, R* bb, R* bc, R* dir) {
  R xba[3], xca[3], xia[3], tmp[3];
  sub3(xb, xa, xba);
  sub3(xc, xa, xca);
  sub3(xi, xa, xia);
  cross3(xba, xca, dir);
  R nn = dot3(dir, dir);
  cross3(xia, xca, tmp);
  R w_iaca = dot3(dir, tmp);
  cross3(xba, xia, tmp);
  R w_baia = dot3(dir, tmp);
  if (nn > (R)1e-24 && w_iaca >= 0 && w_baia >= 0 && nn - w_iaca - w_baia >= 0) {
    *bb = w_iaca / nn;
    *bc = w_baia / nn;
    *ba = (R)1 - *bb - *bc;
  } else {
    bool degen = !(nn > (R)1e-24);
    R best = (R)999999999, r, d, N[3];
    *ba = 1; *bb = 0; *bc = 0;
    if ((degen || nn - w_iaca - w_baia < 0) && (d = vertex_edge_sqd(xi, xb, xc, &r, N)) < best) { best = d; *bb = (R)1 - r; *bc = r; *ba = 0; }
    if ((degen || w_iaca < 0) && (d = vertex_edge_sqd(xi, xa, xc, &r, N)) < best) { best = d; *bb = 0; *bc = r; *ba = (R)1 - *bb - *bc; }
    if ((degen || w_baia < 0) && (d = vertex_edge_sqd(xi, xa, xb, &r, N)) < best) { best = d; *bb = r; *bc = 0; *ba = (R)1 - *bb - *bc; }
  }
  for (int k = 0; k < 3; k++) dir[k] = xi[k] - xa[k] * *ba - xb[k] * *bb - xc[k] * *bc;
  return dot3(dir, dir);
}

// point_triangle_ccd_broadphase / edge_edge_ccd_broadphase (PdIpcGpuFunc.cu:2346-2458)
template <class R>
static bool pt_ccd_filter(const R* p, const R* t0, const R* t1, const R* t2, const R* dp, const R* d0, const R* d1, const R* d2, R dist) {
  for (int k = 0; k < 3; k++) {
    R tmin = min3(t0[k], t1[k], t2[k]);
    tmin = min3(tmin, t0[k] + d0[k], t1[k] + d1[k]);
    tmin = min2(tmin, t2[k] + d2[k]);
    R tmax = max3(t0[k], t1[k], t2[k]);
    tmax = max3(tmax, t0[k] + d0[k], t1[k] + d1[k]);
    tmax = max2(tmax, t2[k] + d2[k]);
    R pmin = min2(p[k], p[k] + dp[k]), pmax = max2(p[k], p[k] + dp[k]);
    if ((pmin - tmax) > dist || (tmin - pmax) > dist) return false;
  }
  return true;
}
template <class R>
static bool ee_ccd_filter(const R* a0, const R* a1, const R* b0, const R* b1, const R* da0, const R* da1, const R* db0, const R* db1, R dist) {
  for (int k = 0; k < 3; k++) {
    R mina = min2(a0[k], a0[k] + da0[k]); mina = min2(mina, a1[k]); mina = min2(mina, a1[k] + da1[k]);
    R maxa = max2(a0[k], a0[k] + da0[k]); maxa = max2(maxa, a1[k]); maxa = max2(maxa, a1[k] + da1[k]);
    R minb = min2(b0[k], b0[k] + db0[k]); minb = min2(minb, b1[k]); minb = min2(minb, b1[k] + db1[k]);
    R maxb = max2(b0[k], b0[k] + db0[k]); maxb = max2(maxb, b1[k]); maxb = max2(maxb, b1[k] + db1[k]);
    if ((mina - maxb) > dist || (minb - maxa) > dist) return false;
  }
  return true;
}
// point_triangle_cd_broadphase / edge_edge_cd_broadphase (PdIpcGpuFunc.cu:2327-2423)
template <class R>
static bool pt_cd_filter(const R* p, const R* t0, const R* t1, const R* t2, R dist) {
  for (int k = 0; k < 3; k++) {
    R tmin = min3(t0[k], t1[k], t2[k]), tmax = max3(t0[k], t1[k], t2[k]);
    if ((p[k] - tmax) > dist || (tmin - p[k]) > dist) return false;
  }
  return true;
}
template <class R>
static bool ee_cd_filter(const R* a0, const R* a1, const R* b0, const R* b1, R dist) {
  for (int k = 0; k < 3; k++) {
    R mina = min2(a0[k], a1[k]), maxa = max2(a0[k], a1[k]);
    R minb = min2(b0[k], b1[k]), maxb = max2(b0[k], b1[k]);
    if ((mina - maxb) > dist || (minb - maxa) > dist) return false;
  }
  return true;
}

// ACCD of ccdVertexTriangleToi / ccdEdgeEdgeToi (PdIpcGpuFunc.cu:2460-2607),
// with D4 under touch_policy = 1.  Pinned by conservativeness against exact
// cubic-root TOIs (tests/test_oracle_math.py) and by the value of head-on
// approaches, a partial geometric series in closed form, for the reference
// constants and other step / gap / iteration settings, with mutation tests
// (tests/test_oracle_closed_forms.py, tests/test_oracle_mutants.py).
template <class R>
static R ccd_toi(int type, R x[4][3], R dx[4][3], int iters, R step, R gapf, int touch_policy) {
  R mov[3];
  for (int k = 0; k < 3; k++) mov[k] = (dx[0][k] + dx[1][k] + dx[2][k] + dx[3][k]) / (R)4.0;
  for (int i = 0; i < 4; i++)
    for (int k = 0; k < 3; k++) dx[i][k] -= mov[k];
  R maxdisp;
  if (type == 0) {
    R m2 = max2(dot3(dx[1], dx[1]), dot3(dx[2], dx[2]));
    m2 = max2(m2, dot3(dx[3], dx[3]));
    maxdisp = sqrt(dot3(dx[0], dx[0])) + sqrt(m2);
  } else {
    R m0 = max2(dot3(dx[0], dx[0]), dot3(dx[1], dx[1]));
    R m1 = max2(dot3(dx[2], dx[2]), dot3(dx[3], dx[3]));
    maxdisp = sqrt(m0) + sqrt(m1);
  }
  if (maxdisp == 0) return 1;
  auto dfun = [&](R* dist2out) -> R {  // returns dFunc; *dist2out = dist2 used for sqrt
    R a, b, c, dir[3];
    if (type == 0) {
      R d2 = vertex_triangle_sqd(x[0], x[1], x[2], x[3], &a, &b, &c, dir);
      *dist2out = d2;
      return d2;
    }
    R d2 = edge_edge_sqd(x[0], x[1], x[2], x[3], &a, &b, dir);
    if (d2 <= 0) {
      R e[3];
      sub3(x[0], x[2], e); d2 = dot3(e, e);
      sub3(x[0], x[3], e); d2 = min2(dot3(e, e), d2);
      sub3(x[1], x[2], e); d2 = min2(dot3(e, e), d2);
      sub3(x[1], x[3], e); d2 = min2(dot3(e, e), d2);
    }
    *dist2out = d2;
    return d2;
  };
  R d2;
  R dF = dfun(&d2);
  R dcur = sqrt(d2);
  if (touch_policy == 1 && !(dcur > 0)) return 0;
  R gap = gapf * dF / dcur;  // = gapf * dcur
  R toc = 0;
  if (ORC_MUT(3)) step = step * (R)0.8 / (R)0.9;
  for (int i = 0; i < iters; i++) {
    R tlb = step * dF / (dcur * maxdisp);
    for (int j = 0; j < 4; j++)
      for (int k = 0; k < 3; k++) x[j][k] += tlb * dx[j][k];
    dF = dfun(&d2);
    dcur = sqrt(d2);
    if (touch_policy == 1 && !(dcur > 0)) return toc;
    if (ORC_MUT(4)) {
      toc += tlb;
      if ((dF / dcur) < gap) break;
    } else {
      if (toc != 0 && (dF / dcur) < gap) break;
      toc += tlb;
    }
    if (toc > 1) { toc = 1; break; }
  }
  return toc;
}

// barrier (PdIpcGpuFunc.cu:2745-2754; Bow/Math/Barrier.h:5-14).  Pinned by closed form.
template <class R>
static R barrier_value(R d2, R dh2, R kappa) {
  R e = 0;
  if (d2 < dh2) {
    R t2 = d2 - dh2;
    e = -kappa * t2 * t2 * log(d2 / dh2) / (dh2 * dh2);
  }
  return e;
}

// ---------------------------------------------------------------------------
// the simulator
// ---------------------------------------------------------------------------
struct Counters {
  int outer = 0, lg = 0, patchPairs = 0, candidates = 0, activePairs = 0;
  double toiMin = 1, residual = 0;
  long long nnzR1 = 0, nnzR2 = 0, n2paths = 0;
};

template <class R>
struct Sim {
  Params prm;
  std::vector<Model> models;   // simulated, in add order
  std::vector<Model> statics;  // static obstacles
  // per simulated model: per-triangle strain stiffness (empty = the model's s0)
  std::vector<std::vector<R>> triStiffness;
  bool initialized = false;

  int totalPoints = 0, simPoints = 0, clothPoints = 0, nTets = 0;
  std::vector<R> X, Xrest, X0, Xold, Dir, Pre, Next, Vel;
  std::vector<R> mass, invMass, grav;
  std::vector<R> initRhs, inertialDiag, elasticDiag, sysDiag, sysRhs, colDiag, colRhs, firstTerm;
  // cloth (one entry per cloth point, neighbour 0 = itself)
  std::vector<std::vector<int>> cNbr;
  std::vector<std::vector<R>> cStrain, cBend, cRest, cLbo;
  std::vector<R> cH0, cBendArea;
  // tets
  std::vector<int> tets;
  std::vector<R> tetDminv, tetW;
  // position constraints
  std::vector<int> fixFlag, fixIdx;
  std::vector<R> fixK;
  // R1 / R2 (PdIpcSimulator.cpp:633-883)
  std::vector<std::vector<int>> r1Idx;
  std::vector<std::vector<R>> r1Val, r1Op;
  std::vector<std::vector<int>> r2Idx;
  std::vector<std::vector<R>> r2Row;           // M^2 row operators
  std::vector<std::vector<std::vector<int>>> r2PathK;  // per entry: middle index k
  std::vector<std::vector<std::vector<R>>> r2PathA;    // per entry: a_ik * a_kj
  std::vector<std::vector<R>> r2Const;
  std::vector<std::vector<int>> r2FirstR1Slot;         // slot in r1 row or -1
  // collision
  std::vector<std::array<int, 3>> bFaces;  // global ids
  std::vector<std::array<int, 2>> bEdges;
  std::vector<std::vector<int>> patchFaces, patchEdges, patchPoints, patchExPoints;
  std::vector<char> patchStatic;
  std::vector<R> colMass;
  std::vector<std::array<int, 5>> cands;
  std::vector<std::array<int, 2>> patchPairs;
  bool activeCollision = false;
  Counters cnt;

  R inflation() const { return prm.v[P_INFLATION] >= 0 ? (R)prm.v[P_INFLATION] : (R)(0.5 * prm.v[P_DHAT]); }

  // ---------------- init: PdIpcSimulator::initialization (PdIpcSimulator.cpp:34-268)
  int init() {
    const R dt = (R)prm.v[P_DT];
    std::vector<std::vector<int>> posCons;
    std::vector<R> posK;
    std::vector<std::vector<std::array<int, 4>>> tetCons;
    std::vector<R> tetK;
    totalPoints = 0;
    clothPoints = 0;
    // per-model init, simulated models first (PdIpcSimulator.cpp:69-124)
    for (size_t mi = 0; mi < models.size(); mi++) {
      Model& m = models[mi];
      int base = totalPoints;
      if (m.kind == KIND_CLOTH) {
        m.topo = build_surface_topo(m.nv, m.F);
        cloth_register_boundary(m);
        cloth_set_patches(m);
        post_process_boundary(m);
        init_cloth(m, (int)mi, base, posCons, posK);
        clothPoints += m.nv;
      } else {
        tet_register_boundary(m);
        tet_set_patches(m);
        post_process_boundary(m);
        init_tet(m, base, tetCons, tetK, posCons, posK);
      }
      totalPoints += m.nv;
    }
    simPoints = totalPoints;
    for (auto& m : statics) {
      m.topo = build_surface_topo(m.nv, m.F);
      cloth_register_boundary(m);
      cloth_set_patches(m);
      post_process_boundary(m);
      totalPoints += m.nv;
    }
    // points: sim first, static after (PdIpcSimulator.cpp:182-194)
    X.assign(3 * totalPoints, 0);
    {
      int off = 0;
      for (auto* list : {&models, &statics})
        for (auto& m : *list) {
          for (int i = 0; i < 3 * m.nv; i++) X[3 * off + i] = (R)m.V[i];
          off += m.nv;
        }
    }
    Xrest = X;
    Vel.assign(3 * simPoints, 0);
    grav.assign(3 * simPoints, 0);
    for (int i = 0; i < simPoints; i++) grav[3 * i + 1] = (R)prm.v[P_GRAVITY] * mass[i];
    // setupHessMat (PdIpcSimulator.cpp:270-414)
    std::vector<std::map<int, R>> H(simPoints);
    fixFlag.assign(simPoints, -1);
    for (size_t i = 0; i < posCons.size(); i++)
      for (int v : posCons[i]) {
        fixFlag[v] = (int)fixIdx.size();
        fixIdx.push_back(v);
        fixK.push_back(posK[i]);
      }
    for (int i = 0; i < simPoints; i++)
      if (fixFlag[i] != -1) H[i][i] += fixK[fixFlag[i]];
    for (size_t i = 0; i < tetCons.size(); i++)
      for (auto& ele : tetCons[i]) {
        R E[3][3];
        for (int r = 0; r < 3; r++)
          for (int k = 0; k < 3; k++) E[r][k] = X[3 * ele[r + 1] + k] - X[3 * ele[0] + k];
        R det = E[0][0] * (E[1][1] * E[2][2] - E[1][2] * E[2][1]) - E[0][1] * (E[1][0] * E[2][2] - E[1][2] * E[2][0]) +
                E[0][2] * (E[1][0] * E[2][1] - E[1][1] * E[2][0]);
        R inv[3][3];  // E^{-1}
        inv[0][0] = (E[1][1] * E[2][2] - E[1][2] * E[2][1]) / det;
        inv[0][1] = (E[0][2] * E[2][1] - E[0][1] * E[2][2]) / det;
        inv[0][2] = (E[0][1] * E[1][2] - E[0][2] * E[1][1]) / det;
        inv[1][0] = (E[1][2] * E[2][0] - E[1][0] * E[2][2]) / det;
        inv[1][1] = (E[0][0] * E[2][2] - E[0][2] * E[2][0]) / det;
        inv[1][2] = (E[0][2] * E[1][0] - E[0][0] * E[1][2]) / det;
        inv[2][0] = (E[1][0] * E[2][1] - E[1][1] * E[2][0]) / det;
        inv[2][1] = (E[0][1] * E[2][0] - E[0][0] * E[2][1]) / det;
        inv[2][2] = (E[0][0] * E[1][1] - E[0][1] * E[1][0]) / det;
        R vol = fabs(det / (R)6.0);
        R w = tetK[i] * vol;
        for (int k = 0; k < 4; k++) tets.push_back(ele[k]);
        tetW.push_back(w);
        // rest.data() of the column-major inverse == row-major Dm^{-1}
        R idm[9];
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++) idm[3 * r + c] = inv[c][r];
        for (int k = 0; k < 9; k++) tetDminv.push_back(idm[k]);
        R half[3][4];
        for (int r = 0; r < 3; r++) {
          half[r][0] = -idm[r] - idm[3 + r] - idm[6 + r];
          half[r][1] = idm[r];
          half[r][2] = idm[3 + r];
          half[r][3] = idm[6 + r];
        }
        for (int r = 0; r < 4; r++)
          for (int c = 0; c < 4; c++) {
            R s = half[0][r] * half[0][c] + half[1][r] * half[1][c] + half[2][r] * half[2][c];
            H[ele[r]][ele[c]] += s * w;
          }
      }
    nTets = (int)tetW.size();
    for (size_t i = 0; i < cNbr.size(); i++) {
      int pi = cNbr[i][0];
      for (size_t k = 1; k < cNbr[i].size(); k++) {
        R st = cStrain[i][k];
        H[pi][pi] += st;
        H[pi][cNbr[i][k]] += -st;
      }
    }
    for (size_t i = 0; i < cNbr.size(); i++) {
      size_t L = cLbo[i].size();
      for (size_t r = 0; r < L; r++)
        for (size_t c = 0; c < L; c++) H[cNbr[i][r]][cNbr[i][c]] += cBendArea[i] * (cLbo[i][r] * cLbo[i][c]);
    }
    elasticDiag.assign(simPoints, 0);
    for (int i = 0; i < simPoints; i++) elasticDiag[i] = H[i].count(i) ? H[i][i] : (R)0;
    for (int i = 0; i < simPoints; i++) H[i][i] += mass[i] / (dt * dt);
    build_solver(H);
    init_collision();
    // buffers
    X0 = X; Xold = X; Pre = X; Next = X; Dir.assign(3 * totalPoints, 0);
    initRhs.assign(3 * simPoints, 0);
    inertialDiag.assign(simPoints, 0);
    sysDiag.assign(simPoints, 0);
    sysRhs.assign(3 * simPoints, 0);
    firstTerm.assign(3 * simPoints, 0);
    colDiag.assign(totalPoints, 0);
    colRhs.assign(3 * totalPoints, 0);
    initialized = true;
    return 0;
  }

  // PdIpcClothModel::init (PdIpcClothModel.h:142-343)
  void init_cloth(Model& m, int modelIndex, int base, std::vector<std::vector<int>>& posCons, std::vector<R>& posK) {
    int nv = m.nv, nf = (int)m.F.size() / 3, ne = (int)m.topo.edges.size();
    auto P = [&](int i, R* p) { p[0] = (R)m.V[3 * i]; p[1] = (R)m.V[3 * i + 1]; p[2] = (R)m.V[3 * i + 2]; };
    std::vector<R> pArea(nv, 0);
    for (int f = 0; f < nf; f++) {
      R a[3], b[3], c[3], e1[3], e2[3], n[3];
      P(m.F[3 * f], a); P(m.F[3 * f + 1], b); P(m.F[3 * f + 2], c);
      sub3(b, a, e1); sub3(c, a, e2); cross3(e1, e2, n);
      R area = (R)0.5 * fabs(sqrt(dot3(n, n)));
      for (int k = 0; k < 3; k++) pArea[m.F[3 * f + k]] += area / 3;
    }
    for (int v = 0; v < nv; v++) {
      R ms = pArea[v] * (R)m.density;
      mass.push_back(ms);
      invMass.push_back((R)1.0 / ms);
      cBendArea.push_back(pArea[v] * (R)m.s1);
    }
    std::vector<R> cot(ne, 0);
    for (int e = 0; e < ne; e++) {
      auto ed = m.topo.edges[e];
      R p1[3], p2[3];
      P(ed[0], p1); P(ed[1], p2);
      for (int fid : m.topo.edgeFaces[e]) {
        int opp = -1;
        for (int k = 0; k < 3; k++) {
          int fv = m.F[3 * fid + k];
          if (fv != ed[0] && fv != ed[1]) opp = fv;
          if (opp != -1) break;
        }
        R p[3], x10[3], x20[3];
        P(opp, p);
        sub3(p1, p, x10); sub3(p2, p, x20);
        R l10 = sqrt(dot3(x10, x10)), l20 = sqrt(dot3(x20, x20));
        R th = acos(dot3(x10, x20) / (l10 * l20));
        R ct = (R)1.0 / tan(th);
        cot[e] += (R)-0.5 * ct;
      }
    }
    for (int v = 0; v < nv; v++) {
      R pv[3];
      P(v, pv);
      int nb = (int)m.topo.pointEdges[v].size();
      std::vector<int> nbr(nb + 1);
      std::vector<R> lbo(nb + 1, 0), rest(nb + 1, 0), st(nb + 1, (R)m.s0), bd(nb + 1, 0);
      std::vector<R> q[3];
      for (int k = 0; k < 3; k++) { q[k].assign(nb + 1, 0); q[k][0] = pv[k]; }
      nbr[0] = v + base;
      R sum = 0;
      for (int k = 0; k < nb; k++) {
        int e = m.topo.pointEdges[v][k];
        auto ed = m.topo.edges[e];
        int ve = v == ed[0] ? ed[1] : ed[0];
        nbr[k + 1] = ve + base;
        R pe[3], d[3];
        P(ve, pe);
        sub3(pv, pe, d);
        rest[k + 1] = sqrt(dot3(d, d));
        if (!triStiffness.empty() && !triStiffness[modelIndex].empty()) {
          const auto& fs = m.topo.edgeFaces[e];
          R acc = 0;
          for (int fid : fs) acc += triStiffness[modelIndex][fid];
          st[k + 1] = acc / (R)(double)fs.size();
        }
        R ew = cot[e] / pArea[v];
        sum += ew;
        lbo[k + 1] = ew;
        for (int j = 0; j < 3; j++) q[j][k + 1] = pe[j];
      }
      lbo[0] = -sum;
      R curv[3] = {0, 0, 0};
      for (int k = 0; k < nb + 1; k++) {
        for (int j = 0; j < 3; j++) curv[j] += lbo[k] * q[j][k];
        bd[k] = lbo[k] * (R)m.s1 * pArea[v];
      }
      cNbr.push_back(nbr);
      cStrain.push_back(st);
      cBend.push_back(bd);
      cRest.push_back(rest);
      cLbo.push_back(lbo);
      cH0.push_back(sqrt(dot3(curv, curv)));
    }
    std::vector<int> pc;
    for (int f : m.fixedIds) pc.push_back(f + base);
    posCons.push_back(pc);
    posK.push_back((R)m.s2);
  }

  // PdIpcVolumetricModel::init (PdIpcVolumetricModel.h:183-273); node volume is the
  // plain sum of adjacent tet volumes (Model/VolumetricMesh/BaseTetMeshHandle.cpp:105-124).
  void init_tet(Model& m, int base, std::vector<std::vector<std::array<int, 4>>>& tetCons, std::vector<R>& tetK,
                std::vector<std::vector<int>>& posCons, std::vector<R>& posK) {
    int nt = (int)m.T.size() / 4;
    std::vector<R> vol(m.nv, 0);
    std::vector<std::array<int, 4>> list(nt);
    for (int e = 0; e < nt; e++) {
      R E[3][3];
      for (int r = 0; r < 3; r++)
        for (int k = 0; k < 3; k++) E[r][k] = (R)m.V[3 * m.T[4 * e + r + 1] + k] - (R)m.V[3 * m.T[4 * e] + k];
      R det = E[0][0] * (E[1][1] * E[2][2] - E[1][2] * E[2][1]) - E[0][1] * (E[1][0] * E[2][2] - E[1][2] * E[2][0]) +
              E[0][2] * (E[1][0] * E[2][1] - E[1][1] * E[2][0]);
      R tv = fabs(det) / (R)6.0;
      for (int k = 0; k < 4; k++) vol[m.T[4 * e + k]] += tv;
      list[e] = {m.T[4 * e] + base, m.T[4 * e + 1] + base, m.T[4 * e + 2] + base, m.T[4 * e + 3] + base};
    }
    for (int i = 0; i < m.nv; i++) {
      mass.push_back(vol[i] * (R)m.density);
      invMass.push_back((R)1.0 / (vol[i] * (R)m.density));
    }
    tetCons.push_back(list);
    tetK.push_back((R)m.s0);
    std::vector<int> pc;
    for (int f : m.fixedIds) pc.push_back(f + base);
    posCons.push_back(pc);
    posK.push_back((R)m.s2);
  }

  // gpuSolverInit R1/R2 part (PdIpcSimulator.cpp:633-883).  The R2 pattern is the
  // value pattern of R1*R1 (|v| >= 1e-12), as Eigen's product followed by the
  // drop loop yields.  Pinned: dense closed form of the Chebyshev map on tiny systems.
  void build_solver(const std::vector<std::map<int, R>>& H) {
    const R IG = (R)1e-12;
    int n = simPoints;
    r1Idx.assign(n, {});
    r1Val.assign(n, {});
    for (int i = 0; i < n; i++) {
      r1Idx[i].push_back(i);
      r1Val[i].push_back(0);
    }
    // Eigen column-major iteration: for each column k, rows ascending; H is symmetric
    // so row i's off-diagonal columns arrive in ascending order.
    for (int i = 0; i < n; i++)
      for (auto& kv : H[i]) {
        if (fabs(kv.second) < IG) continue;
        if (kv.first != i) {
          r1Idx[i].push_back(kv.first);
          r1Val[i].push_back(-kv.second);
        }
      }
    r1Op = r1Val;
    cnt.nnzR1 = 0;
    for (auto& r : r1Idx) cnt.nnzR1 += (long long)r.size();
    // R2 = R1 * R1 values
    r2Idx.assign(n, {});
    r2PathK.assign(n, {});
    r2PathA.assign(n, {});
    r2Const.assign(n, {});
    r2FirstR1Slot.assign(n, {});
    r2Row.assign(n, {});
    std::vector<R> acc(n, 0);
    std::vector<char> seen(n, 0);
    std::vector<int> touched;
    cnt.nnzR2 = 0;
    cnt.n2paths = 0;
    for (int i = 0; i < n; i++) {
      touched.clear();
      // structural product including the explicit zero diagonal of R1
      for (size_t a = 0; a < r1Idx[i].size(); a++) {
        int k = r1Idx[i][a];
        R vik = r1Val[i][a];
        for (size_t b = 0; b < r1Idx[k].size(); b++) {
          int j = r1Idx[k][b];
          if (!seen[j]) { seen[j] = 1; touched.push_back(j); acc[j] = 0; }
          acc[j] += vik * r1Val[k][b];
        }
      }
      std::sort(touched.begin(), touched.end());
      for (int j : touched) {
        R value = acc[j];
        seen[j] = 0;
        if (fabs(value) < IG) continue;
        // 2-path list for (i, j) (PdIpcSimulator.cpp:732-765)
        std::vector<int> ks;
        std::vector<R> as;
        for (size_t r = 0; r < r1Idx[i].size(); r++) {
          int rid = r1Idx[i][r];
          for (size_t c = 0; c < r1Idx[j].size(); c++) {
            if (rid != r1Idx[j][c]) continue;
            R ajk = r1Val[i][r] * r1Val[j][c];
            if (fabs(ajk) < IG) continue;
            ks.push_back(rid);
            as.push_back(ajk);
          }
        }
        if (ks.empty()) continue;  // hessR2IndexValid = false
        r2Idx[i].push_back(j);
        r2PathK[i].push_back(ks);
        r2PathA[i].push_back(as);
        r2Const[i].push_back(i == j ? (R)1 : (R)0);
        int slot = -1;
        for (size_t k = 0; k < r1Idx[i].size(); k++)
          if (r1Idx[i][k] == j) { slot = (int)k; break; }
        r2FirstR1Slot[i].push_back(slot);
        r2Row[i].push_back(0);
        cnt.n2paths += (long long)ks.size();
      }
      cnt.nnzR2 += (long long)r2Idx[i].size();
    }
  }

  // gpuCollisionInit (PdIpcSimulator.cpp:1151-1504) minus the BVH (the pair set is
  // computed by sweep-and-prune, which finds exactly the overlapping pairs).
  void init_collision() {
    int pointOff = 0, faceOff = 0, edgeOff = 0;
    std::vector<std::vector<int>> fEdges, fPoints;
    std::vector<int> modelOfPatch;
    int mi = 0;
    for (auto* list : {&models, &statics}) {
      bool isStatic = list == &statics;
      for (auto& m : *list) {
        for (auto& f : m.bFaces) bFaces.push_back({f[0] + pointOff, f[1] + pointOff, f[2] + pointOff});
        for (auto& e : m.bEdges) bEdges.push_back({e[0] + pointOff, e[1] + pointOff});
        for (auto& p : m.patches) {
          std::vector<int> t;
          for (int f : p) t.push_back(f + faceOff);
          if (!t.empty()) {
            patchFaces.push_back(t);
            patchStatic.push_back(isStatic ? 1 : 0);
          }
        }
        for (size_t j = 0; j < m.bFaces.size(); j++) {
          std::vector<int> te, tp;
          for (int e : m.facesSharedEdges[j]) te.push_back(e + edgeOff);
          for (int p : m.facesSharedPoints[j]) tp.push_back(p + pointOff);
          fEdges.push_back(te);
          fPoints.push_back(tp);
        }
        pointOff += m.nv;
        faceOff += (int)m.bFaces.size();
        edgeOff += (int)m.bEdges.size();
        mi++;
      }
    }
    int P = (int)patchFaces.size();
    patchEdges.assign(P, {});
    patchPoints.assign(P, {});
    patchExPoints.assign(P, {});
    for (int i = 0; i < P; i++) {
      std::set<int> es, ps, xs;
      for (int fid : patchFaces[i]) {
        for (int e : fEdges[fid]) es.insert(e);
        for (int p : fPoints[fid]) ps.insert(p);
        for (int k = 0; k < 3; k++) xs.insert(bFaces[fid][k]);
      }
      patchEdges[i].assign(es.begin(), es.end());
      patchPoints[i].assign(ps.begin(), ps.end());
      patchExPoints[i].assign(xs.begin(), xs.end());
    }
    colMass.assign(totalPoints, ORC_MUT(7) ? (R)1 : (R)prm.v[P_STATIC_MASS]);
    for (int i = 0; i < simPoints; i++) colMass[i] = mass[i];
  }

  // ---------------- step: PdIpcSimulator::advance (PdIpcSimulator.cpp:1683-1796)
  void predictive(R dt) {  // predictiveKernel (PdIpcGpuFunc.cu:88-147)
    for (int t = 0; t < simPoints; t++) {
      if (fixFlag[t] != -1) {
        Vel[3 * t] = Vel[3 * t + 1] = Vel[3 * t + 2] = 0;
        continue;
      }
      R m = invMass[t] * dt;
      R v[3];
      for (int k = 0; k < 3; k++) {
        v[k] = Vel[3 * t + k];
        v[k] += m * (grav[3 * t + k] + (R)0 + (R)0);
        v[k] *= dt;
        v[k] += X[3 * t + k];
      }
      m = (R)1.0 / (dt * m);
      inertialDiag[t] = (R)1.0 / (invMass[t] * dt * dt);
      for (int k = 0; k < 3; k++) {
        initRhs[3 * t + k] = m * v[k];
        X[3 * t + k] = v[k];
      }
    }
  }

  // solveVolumetricStrainConstraintKernel (PdIpcGpuFunc.cu:439-599)
  void local_tets(const std::vector<R>& x, std::vector<R>& rhs) {
    const R lo = (R)prm.v[P_SIG_LO], hi = (R)prm.v[P_SIG_HI];
    for (int t = 0; t < nTets; t++) {
      const int* ti = &tets[4 * t];
      const R* idm = &tetDminv[9 * t];
      R Ds[9];
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) Ds[3 * r + c] = x[3 * ti[c + 1] + r] - x[3 * ti[0] + r];
      R F[9];
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) F[3 * r + c] = Ds[3 * r] * idm[c] + Ds[3 * r + 1] * idm[3 + c] + Ds[3 * r + 2] * idm[6 + c];
      R Pm[9];
      strain_projection<R>(F, lo, hi, Pm);
      R half[3][4];
      for (int r = 0; r < 3; r++) {
        half[r][0] = -idm[r] - idm[3 + r] - idm[6 + r];
        half[r][1] = idm[r];
        half[r][2] = idm[3 + r];
        half[r][3] = idm[6 + r];
      }
      R w = tetW[t];
      for (int c = 0; c < 4; c++)
        for (int r = 0; r < 3; r++) {
          R v = Pm[3 * r] * half[0][c] + Pm[3 * r + 1] * half[1][c] + Pm[3 * r + 2] * half[2][c];
          rhs[3 * ti[c] + r] += v * w;
        }
    }
  }

  // solveClothPointsStrainAndBendingConstraintKernel (PdIpcGpuFunc.cu:627-728).
  // Pinned at rest (rhs = K_elastic x) and at uniformly scaled and shifted rest
  // shapes of a curved sheet, where springs return l0 * d_hat and the
  // curvature vector is rescaled back to its rest value (closed forms, with
  // mutation tests of both ratios).
  void local_cloth(const std::vector<R>& x, std::vector<R>& rhs) {
    for (int tid = 0; tid < clothPoints; tid++) {
      const auto& nb = cNbr[tid];
      int pid = nb[0];
      R xi[3] = {x[3 * pid], x[3 * pid + 1], x[3 * pid + 2]};
      R tx[3] = {0, 0, 0};
      R lbo = cLbo[tid][0];
      R aq[3] = {xi[0] * lbo, xi[1] * lbo, xi[2] * lbo};
      for (size_t i = 1; i < nb.size(); i++) {
        const R* nx = &x[3 * nb[i]];
        R d[3] = {xi[0] - nx[0], xi[1] - nx[1], xi[2] - nx[2]};
        R cur = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        R rc = cStrain[tid][i] * (ORC_MUT(5) ? cur / cRest[tid][i] : cRest[tid][i] / cur);
        for (int k = 0; k < 3; k++) tx[k] += rc * d[k];
        lbo = cLbo[tid][i];
        for (int k = 0; k < 3; k++) aq[k] += nx[k] * lbo;
      }
      R an = sqrt(dot3(aq, aq));
      if (an >= (R)1e-10) {
        R c = ORC_MUT(6) ? an / cH0[tid] : cH0[tid] / an;
        for (int k = 0; k < 3; k++) aq[k] *= c;
      }
      R b = cBend[tid][0];
      for (int k = 0; k < 3; k++) tx[k] += b * aq[k];
      for (int k = 0; k < 3; k++) rhs[3 * pid + k] += tx[k];
      for (size_t j = 1; j < nb.size(); j++) {
        b = cBend[tid][j];
        for (int k = 0; k < 3; k++) rhs[3 * nb[j] + k] += b * aq[k];
      }
    }
  }

  // doLocalProjection (PdIpcSimulator.cpp:1624-1681)
  void local_projection(const std::vector<R>& x, std::vector<R>& rhs) {
    for (int i = 0; i < 3 * simPoints; i++) rhs[i] = initRhs[i];
    if (nTets > 0) local_tets(x, rhs);
    if (clothPoints > 0) local_cloth(x, rhs);
    if (activeCollision)
      for (int i = 0; i < 3 * simPoints; i++) rhs[i] += colRhs[i];
    for (size_t t = 0; t < fixIdx.size(); t++) {  // solvePointsConstraintKernel (PdIpcGpuFunc.cu:764-790)
      int v = fixIdx[t];
      for (int k = 0; k < 3; k++) rhs[3 * v + k] += fixK[t] * Xrest[3 * v + k];
    }
  }

  // updateR1/R2/R2RowOperatorsKernel (PdIpcGpuFunc.cu:899-974)
  void update_operators() {
    const R w = (R)prm.v[P_OMEGA];
#pragma omp parallel for schedule(static)
    for (int i = 0; i < simPoints; i++)
      for (size_t a = 0; a < r1Idx[i].size(); a++) r1Op[i][a] = r1Val[i][a] / sysDiag[i];
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < simPoints; i++)
      for (size_t e = 0; e < r2Idx[i].size(); e++) {
        R val = 0;
        for (size_t p = 0; p < r2PathK[i][e].size(); p++) val += r2PathA[i][e][p] / (sysDiag[i] * sysDiag[r2PathK[i][e][p]]);
        val *= w * w;
        val += (1 - w) * (1 - w) * r2Const[i][e];
        int s = r2FirstR1Slot[i][e];
        if (s > -1) val += 2 * (1 - w) * w * r1Op[i][s];
        r2Row[i][e] = val;
      }
  }

  // chebyshevJacobiR2SolverHost (PdIpcGpuFunc.cu:1245-1346) with the kernels at
  // :1041-1074 (first term), :1076-1129 (x1 = M^2 x0 + c), :1185-1242 (Chebyshev).
  void cheb_solve(std::vector<R>*& pre, std::vector<R>*& cur, std::vector<R>*& next) {
    const R w = (R)prm.v[P_OMEGA];
    const int n = simPoints;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++)
      for (int k = 0; k < 3; k++) {
        R val = 0;
        for (size_t a = 1; a < r1Idx[i].size(); a++) {
          int j = r1Idx[i][a];
          val += r1Op[i][a] * (sysRhs[3 * j + k] / sysDiag[j]);
        }
        firstTerm[3 * i + k] = w * w * val + (2 - w) * w * sysRhs[3 * i + k] / sysDiag[i];
      }
    auto spmv_row = [&](const std::vector<R>& xin, int i, int k) {
      R val = 0;
      for (size_t e = 0; e < r2Idx[i].size(); e++) val += r2Row[i][e] * xin[3 * r2Idx[i][e] + k];
      return val;
    };
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++)
      for (int k = 0; k < 3; k++) (*next)[3 * i + k] = spmv_row(*cur, i, k) + firstTerm[3 * i + k];
    std::swap(cur, pre);
    std::swap(cur, next);
    R rho = (R)prm.v[P_RHO];
    rho *= rho;
    R omega = (R)2.0 / ((R)2 - rho);
    int maxIter = (int)prm.v[P_CHEB_ITERS];
    for (int it = 1; it < maxIter; it++) {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
          int idx = 3 * i + k;
          R val = spmv_row(*cur, i, k) + firstTerm[idx];
          (*next)[idx] = omega * (val - (*pre)[idx]) + (*pre)[idx];
        }
      omega = (R)4.0 / ((R)4 - rho * omega);
      std::swap(cur, pre);
      std::swap(cur, next);
    }
  }

  // LocalGlobalIteration (PdIpcSimulator.cpp:1554-1622)
  void local_global() {
    for (int i = 0; i < simPoints; i++) sysDiag[i] = elasticDiag[i] + inertialDiag[i] + (activeCollision ? colDiag[i] : (R)0);
    update_operators();
    int lg = (int)prm.v[P_LG_ITERS];
    for (int it = 0; it < lg; it++) {
      local_projection(X, sysRhs);
      // the three rotating buffers: X is "cur"; after a multiple of 3 writes the
      // final iterate is back in X (the reference relies on 15 % 3 == 0).
      std::vector<R>* pre = &Pre;
      std::vector<R>* cur = &X;
      std::vector<R>* next = &Next;
      cheb_solve(pre, cur, next);
      if (cur != &X) {
        // non-multiple-of-3 sweep counts: copy the final iterate into X
        std::vector<R> tmp = *cur;
        for (int i = 0; i < 3 * simPoints; i++) X[i] = tmp[i];
      }
      cnt.lg++;
    }
  }

  // patch swept boxes (updateCcdPatchBboxesKernel + getFaceBbox, PdIpcGpuFunc.cu:1350-1473)
  void patch_boxes(const std::vector<R>& x, const std::vector<R>& dx, std::vector<std::array<R, 6>>& boxes) {
    R infl = inflation();
    int P = (int)patchFaces.size();
    boxes.assign(P, {});
    for (int p = 0; p < P; p++) {
      std::array<R, 6> b;
      for (int k = 0; k < 3; k++) { b[k] = std::numeric_limits<R>::max(); b[3 + k] = -std::numeric_limits<R>::max(); }
      for (int fid : patchFaces[p]) {
        const auto& f = bFaces[fid];
        for (int k = 0; k < 3; k++) {
          R p0 = x[3 * f[0] + k], p1 = x[3 * f[1] + k], p2 = x[3 * f[2] + k];
          R q0 = p0 + dx[3 * f[0] + k], q1 = p1 + dx[3 * f[1] + k], q2 = p2 + dx[3 * f[2] + k];
          R mn = min3(p0, p1, p2);
          mn = min3(mn, q0, q1);
          mn = min2(mn, q2);
          mn -= infl;
          R mx = max3(p0, p1, p2);
          mx = max3(mx, q0, q1);
          mx = max2(mx, q2);
          mx += infl;
          b[k] = min2(b[k], mn);
          b[3 + k] = max2(b[3 + k], mx);
        }
      }
      boxes[p] = b;
    }
  }

  // gpuCcdCulling (PdIpcSimulator.cpp:1799-1869): overlapping patch pairs p < q
  // (checkOverlap, PdIpcGpuFunc.cu:1615-1626, inclusive) then the record
  // enumeration of getCcdPotentialICPairsListKernel (PdIpcGpuFunc.cu:1838-1948).
  // Pinned: set equality with brute-force overlap (tests).
  void broad_phase(const std::vector<R>& x, const std::vector<R>& dx) {
    std::vector<std::array<R, 6>> boxes;
    patch_boxes(x, dx, boxes);
    int P = (int)boxes.size();
    bool skipStatic = prm.v[P_SKIP_STATIC_PAIRS] != 0;
    // sweep and prune along x
    std::vector<int> order(P);
    for (int i = 0; i < P; i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return boxes[a][0] < boxes[b][0] || (boxes[a][0] == boxes[b][0] && a < b); });
    patchPairs.clear();
    for (int ii = 0; ii < P; ii++) {
      int a = order[ii];
      for (int jj = ii + 1; jj < P; jj++) {
        int b = order[jj];
        if (boxes[b][0] > boxes[a][3]) break;
        const auto& A = boxes[a];
        const auto& B = boxes[b];
        if (A[3] < B[0] || B[3] < A[0]) continue;
        if (A[4] < B[1] || B[4] < A[1]) continue;
        if (A[5] < B[2] || B[5] < A[2]) continue;
        if (skipStatic && patchStatic[a] && patchStatic[b]) continue;
        patchPairs.push_back({min2(a, b), max2(a, b)});
      }
    }
    std::sort(patchPairs.begin(), patchPairs.end());
    cands.clear();
    for (auto& pr : patchPairs) {
      int A = pr[0], B = pr[1];
      for (int p : patchPoints[A])
        for (int f : patchFaces[B]) cands.push_back({0, p, bFaces[f][0], bFaces[f][1], bFaces[f][2]});
      for (int p : patchPoints[B])
        for (int f : patchFaces[A]) cands.push_back({0, p, bFaces[f][0], bFaces[f][1], bFaces[f][2]});
      for (int ea : patchEdges[A])
        for (int eb : patchEdges[B]) cands.push_back({1, bEdges[ea][0], bEdges[ea][1], bEdges[eb][0], bEdges[eb][1]});
    }
    cnt.patchPairs = (int)patchPairs.size();
    cnt.candidates = (int)cands.size();
  }

  // getImpactOfTimeKernel + host min (PdIpcGpuFunc.cu:2610-2742)
  // The minimum over the records: threads take contiguous ranges and keep the
  // first record attaining their strict minimum; merging by (value, record)
  // yields the serial loop's answer (the first record attaining the minimum).
  R compute_toi(const std::vector<R>& x0, const std::vector<R>& x1) {
    R toiMin = 1;
    int64_t best = INT64_MAX;
    int tp = (int)prm.v[P_TOUCH_POLICY];
    const int64_t nc = (int64_t)cands.size();
#pragma omp parallel
    {
      R lmin = 1;
      int64_t lidx = INT64_MAX;
#pragma omp for schedule(static) nowait
      for (int64_t ci = 0; ci < nc; ci++) {
        const auto& c = cands[ci];
        int ids[4] = {c[1], c[2], c[3], c[4]};
        R x[4][3], dx[4][3];
        for (int i = 0; i < 4; i++)
          for (int k = 0; k < 3; k++) {
            x[i][k] = x0[3 * ids[i] + k];
            dx[i][k] = x1[3 * ids[i] + k] - x[i][k];
          }
        R toi = 1;
        if (c[0] == 0) {
          if (ids[0] == ids[1] || ids[0] == ids[2] || ids[0] == ids[3]) continue;
          if (!pt_ccd_filter(x[0], x[1], x[2], x[3], dx[0], dx[1], dx[2], dx[3], (R)prm.v[P_CCD_FILTER])) continue;
        } else {
          if (ids[0] == ids[2] || ids[0] == ids[3] || ids[1] == ids[2] || ids[1] == ids[3]) continue;
          if (!ee_ccd_filter(x[0], x[1], x[2], x[3], dx[0], dx[1], dx[2], dx[3], (R)prm.v[P_CCD_FILTER])) continue;
        }
        toi = ccd_toi<R>(c[0], x, dx, (int)prm.v[P_CCD_ITERS], (R)prm.v[P_CCD_STEP], (R)prm.v[P_CCD_GAP], tp);
        if (toi < lmin) {
          lmin = toi;
          lidx = ci;
        }
      }
#pragma omp critical(orc_toi_min)
      {
        if (lmin < toiMin || (!(toiMin < lmin) && lidx < best)) {
          toiMin = lmin;
          best = lidx;
        }
      }
    }
    if (toiMin < 1) toiMin *= (R)prm.v[P_TOI_SCALE];
    return toiMin;
  }

  // setCollisionConstraintsKernel (PdIpcGpuFunc.cu:2756-2892).  Pinned by
  // closed forms (a point over a static triangle, a point over a simulated
  // triangle of equal mass, edges crossing at right angles: weights, diagonal
  // sums and targets from the geometry) with mutation tests.
  // The activity test (filter and distance) runs over the records in
  // parallel; the accumulation then visits the active records in record order,
  // so the sums are those of the serial loop.
  void collision_constraints(const std::vector<R>& x, R dt) {
    const R dHat = (R)prm.v[P_DHAT], dHat2 = dHat * dHat, kappa = (R)prm.v[P_KAPPA];
    const int64_t nc = (int64_t)cands.size();
    std::vector<uint8_t> act(nc, 0);
#pragma omp parallel for schedule(static)
    for (int64_t ci = 0; ci < nc; ci++) {
      const auto& c = cands[ci];
      int t0 = c[1], t1 = c[2], t2 = c[3], t3 = c[4];
      const R* x0 = &x[3 * t0];
      const R* x1 = &x[3 * t1];
      const R* x2 = &x[3 * t2];
      const R* x3 = &x[3 * t3];
      R dir[3], d2;
      if (c[0] == 0) {
        if (t0 == t1 || t0 == t2 || t0 == t3) continue;
        if (!pt_cd_filter(x0, x1, x2, x3, dHat)) continue;
        R a, b, cc;
        d2 = vertex_triangle_sqd(x0, x1, x2, x3, &a, &b, &cc, dir);
      } else {
        if (t0 == t2 || t0 == t3 || t1 == t2 || t1 == t3) continue;
        if (!ee_cd_filter(x0, x1, x2, x3, dHat)) continue;
        R r, sp;
        d2 = edge_edge_sqd(x0, x1, x2, x3, &r, &sp, dir);
      }
      act[ci] = d2 < dHat2 ? 1 : 0;
    }
    int active = 0;
    for (int64_t ci = 0; ci < nc; ci++) {
      if (!act[ci]) continue;
      const auto& c = cands[ci];
      int t0 = c[1], t1 = c[2], t2 = c[3], t3 = c[4];
      const R* x0 = &x[3 * t0];
      const R* x1 = &x[3 * t1];
      const R* x2 = &x[3 * t2];
      const R* x3 = &x[3 * t3];
      R dir[3], d2, w, mp, mq, rp, rq, dp, dq, dist;
      if (c[0] == 0) {
        if (t0 == t1 || t0 == t2 || t0 == t3) continue;
        if (!pt_cd_filter(x0, x1, x2, x3, dHat)) continue;
        R a, b, cc;
        d2 = vertex_triangle_sqd(x0, x1, x2, x3, &a, &b, &cc, dir);
        if (d2 >= dHat2) continue;
        dist = sqrt(d2);
        mp = colMass[t0];
        mq = a * colMass[t1] + b * colMass[t2] + cc * colMass[t3];
        w = barrier_value(d2, dHat2, kappa) + min2(mp, mq) / (dt * dt);
        colDiag[t0] += w; colDiag[t1] += w; colDiag[t2] += w; colDiag[t3] += w;
        rp = mq / (mp + mq);
        rq = mp / (mp + mq);
        if (ORC_MUT(1)) std::swap(rp, rq);
        dist = dHat - dist;
        dp = dist * rp;
        dq = dist * rq;
        R len = sqrt(dot3(dir, dir));
        for (int k = 0; k < 3; k++) dir[k] /= len;
        R wt[3] = {1, 1, 1};
        if (ORC_MUT(2)) { wt[0] = a; wt[1] = b; wt[2] = cc; }
        for (int k = 0; k < 3; k++) {
          colRhs[3 * t0 + k] += w * (x0[k] + dp * dir[k]);
          colRhs[3 * t1 + k] += w * (x1[k] - wt[0] * dq * dir[k]);
          colRhs[3 * t2 + k] += w * (x2[k] - wt[1] * dq * dir[k]);
          colRhs[3 * t3 + k] += w * (x3[k] - wt[2] * dq * dir[k]);
        }
      } else {
        if (t0 == t2 || t0 == t3 || t1 == t2 || t1 == t3) continue;
        if (!ee_cd_filter(x0, x1, x2, x3, dHat)) continue;
        R r, s;
        d2 = edge_edge_sqd(x0, x1, x2, x3, &r, &s, dir);
        if (d2 >= dHat2) continue;
        dist = sqrt(d2);
        mp = (1 - r) * colMass[t0] + r * colMass[t1];
        mq = (1 - s) * colMass[t2] + s * colMass[t3];
        w = barrier_value(d2, dHat2, kappa) + min2(mp, mq) / (dt * dt);
        colDiag[t0] += w; colDiag[t1] += w; colDiag[t2] += w; colDiag[t3] += w;
        rp = mq / (mp + mq);
        rq = mp / (mp + mq);
        if (ORC_MUT(1)) std::swap(rp, rq);
        R len = sqrt(dot3(dir, dir));
        for (int k = 0; k < 3; k++) dir[k] /= len;
        dist = dHat - dist;
        dp = dist * rp;
        dq = dist * rq;
        for (int k = 0; k < 3; k++) {
          colRhs[3 * t0 + k] += w * (x0[k] + dp * dir[k]);
          colRhs[3 * t1 + k] += w * (x1[k] + dp * dir[k]);
          colRhs[3 * t2 + k] += w * (x2[k] - dq * dir[k]);
          colRhs[3 * t3 + k] += w * (x3[k] - dq * dir[k]);
        }
      }
      active++;
    }
    cnt.activePairs = active;
  }

  void advance() {
    const R dt = (R)prm.v[P_DT];
    const int maxOuter = (int)prm.v[P_MAX_OUTER];
    activeCollision = false;
    cnt.lg = 0;
    cnt.toiMin = 1;
    for (int i = 0; i < 3 * simPoints; i++) { X0[i] = X[i]; Xold[i] = X[i]; }
    predictive(dt);
    int outer = 0;
    for (; outer < maxOuter; outer++) {
      local_global();
      for (int i = 0; i < 3 * simPoints; i++) Dir[i] = X[i] - Xold[i];
      if (outer == 0) broad_phase(Xold, Dir);
      R toi = compute_toi(Xold, X);
      if (toi < cnt.toiMin) cnt.toiMin = (double)toi;
      for (int i = 0; i < 3 * simPoints; i++) X[i] = Xold[i] + toi * Dir[i];
      for (int i = 0; i < simPoints; i++) colDiag[i] = 0;
      for (int i = 0; i < 3 * simPoints; i++) colRhs[i] = 0;
      collision_constraints(X, dt);
      activeCollision = true;
      R res = 0;
      for (int i = 0; i < 3 * simPoints; i++) res += Dir[i] * Dir[i];
      cnt.residual = (double)res;
      if (res < (R)prm.v[P_RESIDUAL_TOL] && outer > 0) break;
      for (int i = 0; i < 3 * simPoints; i++) Xold[i] = X[i];
    }
    cnt.outer = outer < maxOuter ? outer + 1 : maxOuter;
    // updateVelocityKernel (PdIpcGpuFunc.cu:262-280) with damping 1 - dt (PdIpcSimulator.cpp:1789)
    R invt = (R)1.0 / dt * ((R)1. - dt);
    for (int i = 0; i < 3 * simPoints; i++) Vel[i] = (X[i] - X0[i]) * invt;
  }
};

}  // namespace orc

// ---------------------------------------------------------------------------
// C ABI for the test harness (oracle/oracle.py)
// ---------------------------------------------------------------------------
using namespace orc;

// mode 0: FP64 oracle; 1: FP32 mimic; 2: FP64 with kDualDirs forward-mode tangents
struct OracleHandle {
  int useFloat = 0;
  Sim<double> d;
  Sim<float> f;
  Sim<DualD> z;
  std::string err;
};

template <class F>
static auto with_sim(OracleHandle* h, F&& fn) {
  if (h->useFloat == 1) return fn(h->f);
  if (h->useFloat == 2) return fn(h->z);
  return fn(h->d);
}

extern "C" {
// OpenMP threads of the parallel loops (solve, operator update, TOI, barrier
// activity); n <= 0: all processors.  Results do not depend on the count.
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : omp_get_num_procs()); }
int orc_get_threads() { return omp_get_max_threads(); }


void* orc_create(const double* params, int nparams, int useFloat) {
  auto* h = new OracleHandle();
  h->useFloat = useFloat;
  Params p;
  for (int i = 0; i < nparams && i < P_COUNT; i++) p.v[i] = params[i];
  h->d.prm = p;
  h->f.prm = p;
  h->z.prm = p;
  return h;
}

void orc_destroy(void* hp) { delete (OracleHandle*)hp; }

static Model make_model(int kind, const double* v, int nv, const int* el, int nel, int per, double s0, double s1, double s2,
                        double density, int patchDim, const int* fixed, int nfixed) {
  Model m;
  m.kind = kind;
  m.nv = nv;
  m.V.assign(v, v + 3 * nv);
  if (kind == KIND_TET) m.T.assign(el, el + per * nel);
  else m.F.assign(el, el + per * nel);
  m.s0 = s0; m.s1 = s1; m.s2 = s2; m.density = density; m.patchDim = patchDim;
  if (fixed && nfixed > 0) m.fixedIds.assign(fixed, fixed + nfixed);
  return m;
}

int orc_add_cloth(void* hp, const double* v, int nv, const int* f, int nf, double s0, double s1, double s2, double density,
                  int patchDim, const int* fixed, int nfixed) {
  auto* h = (OracleHandle*)hp;
  Model m = make_model(KIND_CLOTH, v, nv, f, nf, 3, s0, s1, s2, density, patchDim, fixed, nfixed);
  h->d.models.push_back(m);
  h->f.models.push_back(m);
  h->z.models.push_back(m);
  return 0;
}

int orc_add_tet(void* hp, const double* v, int nv, const int* t, int nt, double s0, double s1, double s2, double density,
                int patchDim, const int* fixed, int nfixed) {
  auto* h = (OracleHandle*)hp;
  Model m = make_model(KIND_TET, v, nv, t, nt, 4, s0, s1, s2, density, patchDim, fixed, nfixed);
  h->d.models.push_back(m);
  h->f.models.push_back(m);
  h->z.models.push_back(m);
  return 0;
}

int orc_add_static(void* hp, const double* v, int nv, const int* f, int nf, int patchDim) {
  auto* h = (OracleHandle*)hp;
  Model m = make_model(KIND_STATIC, v, nv, f, nf, 3, 0, 0, 0, 1e5, patchDim, nullptr, 0);
  h->d.statics.push_back(m);
  h->f.statics.push_back(m);
  h->z.statics.push_back(m);
  return 0;
}

// per-triangle strain stiffness of simulated model `mi` (a cloth), before
// orc_init; `tangents` (kDualDirs x ntri, row per direction) seeds the dual mode
int orc_set_tri_stiffness(void* hp, int mi, const double* k, int ntri, const double* tangents) {
  auto* h = (OracleHandle*)hp;
  if (mi < 0 || mi >= (int)h->d.models.size() || h->d.models[mi].kind != KIND_CLOTH) return -1;
  if (ntri * 3 != (int)h->d.models[mi].F.size()) return -2;
  auto put = [&](auto& s) {
    using R = typename std::remove_reference<decltype(s.X)>::type::value_type;
    s.triStiffness.resize(s.models.size());
    s.triStiffness[mi].assign(ntri, (R)0);
    for (int t = 0; t < ntri; t++) s.triStiffness[mi][t] = (R)k[t];
  };
  put(h->d);
  put(h->f);
  put(h->z);
  if (tangents)
    for (int t = 0; t < ntri; t++)
      for (int q = 0; q < kDualDirs; q++) h->z.triStiffness[mi][t].d[q] = tangents[(size_t)q * ntri + t];
  return 0;
}

int orc_dual_dirs(void) { return kDualDirs; }

// dual mode: values and tangents (kDualDirs x 3*sim, row per direction) of x or v
int orc_dual_set(void* hp, int which, const double* val, const double* tangents) {
  auto* h = (OracleHandle*)hp;
  auto& s = h->z;
  std::vector<DualD>& a = which == 0 ? s.X : s.Vel;
  int n = 3 * s.simPoints;
  for (int i = 0; i < n; i++) {
    a[i] = DualD(val[i]);
    if (tangents)
      for (int q = 0; q < kDualDirs; q++) a[i].d[q] = tangents[(size_t)q * n + i];
  }
  return 0;
}
int orc_dual_get(void* hp, int which, double* val, double* tangents) {
  auto* h = (OracleHandle*)hp;
  auto& s = h->z;
  const std::vector<DualD>& a = which == 0 ? s.X : s.Vel;
  int n = 3 * s.simPoints;
  for (int i = 0; i < n; i++) {
    val[i] = a[i].v;
    if (tangents)
      for (int q = 0; q < kDualDirs; q++) tangents[(size_t)q * n + i] = a[i].d[q];
  }
  return 0;
}

int orc_init(void* hp) {
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [](auto& s) { return s.init(); });
}

int orc_step(void* hp, int n) {
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [n](auto& s) {
    for (int i = 0; i < n; i++) s.advance();
    return 0;
  });
}

int orc_sizes(void* hp, int* out) {  // total, sim, cloth, tets, patches, boundary faces, boundary edges
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [out](auto& s) {
    out[0] = s.totalPoints; out[1] = s.simPoints; out[2] = s.clothPoints; out[3] = s.nTets;
    out[4] = (int)s.patchFaces.size(); out[5] = (int)s.bFaces.size(); out[6] = (int)s.bEdges.size();
    return 0;
  });
}

int orc_get_x(void* hp, double* out) {
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [out](auto& s) { for (size_t i = 0; i < s.X.size(); i++) out[i] = (double)s.X[i]; return 0; });
}
int orc_set_x(void* hp, const double* in) {  // simulated points only
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [in](auto& s) { for (int i = 0; i < 3 * s.simPoints; i++) s.X[i] = in[i]; return 0; });
}
int orc_get_v(void* hp, double* out) {
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [out](auto& s) { for (size_t i = 0; i < s.Vel.size(); i++) out[i] = (double)s.Vel[i]; return 0; });
}
int orc_set_v(void* hp, const double* in) {
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [in](auto& s) { for (size_t i = 0; i < s.Vel.size(); i++) s.Vel[i] = in[i]; return 0; });
}

// counters: outer, lg, patchPairs, candidates, activePairs, toiMin, residual, nnzR1, nnzR2, n2paths
int orc_counters(void* hp, double* out) {
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [out](auto& s) {
    auto& c = s.cnt;
    out[0] = c.outer; out[1] = c.lg; out[2] = c.patchPairs; out[3] = c.candidates; out[4] = c.activePairs;
    out[5] = c.toiMin; out[6] = c.residual; out[7] = (double)c.nnzR1; out[8] = (double)c.nnzR2; out[9] = (double)c.n2paths;
    return 0;
  });
}

// candidate records of the last broad phase, 5 ints each; returns the count
int orc_candidates(void* hp, int* out, int cap) {
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [out, cap](auto& s) {
    int n = (int)s.cands.size();
    if (out)
      for (int i = 0; i < n && i < cap; i++)
        for (int k = 0; k < 5; k++) out[5 * i + k] = s.cands[i][k];
    return n;
  });
}

// patch pairs of the last broad phase (2 ints each); returns the count
int orc_patch_pairs(void* hp, int* out, int cap) {
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [out, cap](auto& s) {
    int n = (int)s.patchPairs.size();
    if (out)
      for (int i = 0; i < n && i < cap; i++) { out[2 * i] = s.patchPairs[i][0]; out[2 * i + 1] = s.patchPairs[i][1]; }
    return n;
  });
}

// ---- teacher-forced phase entry points (x arrays cover all points, 3*total) ----
int orc_debug_broad_phase(void* hp, const double* xold, const double* dx) {
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [&](auto& s) {
    using R = typename std::remove_reference<decltype(s.X[0])>::type;
    std::vector<R> a(xold, xold + 3 * s.totalPoints), b(dx, dx + 3 * s.totalPoints);
    s.broad_phase(a, b);
    return (int)s.cands.size();
  });
}
double orc_debug_toi(void* hp, const double* x0, const double* x1) {
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [&](auto& s) {
    using R = typename std::remove_reference<decltype(s.X[0])>::type;
    std::vector<R> a(x0, x0 + 3 * s.totalPoints), b(x1, x1 + 3 * s.totalPoints);
    return (double)s.compute_toi(a, b);
  });
}
int orc_debug_collision(void* hp, const double* x, double* diag, double* rhs) {
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [&](auto& s) {
    using R = typename std::remove_reference<decltype(s.X[0])>::type;
    std::vector<R> a(x, x + 3 * s.totalPoints);
    std::fill(s.colDiag.begin(), s.colDiag.end(), (R)0);
    std::fill(s.colRhs.begin(), s.colRhs.end(), (R)0);
    s.collision_constraints(a, (R)s.prm.v[P_DT]);
    for (int i = 0; i < s.simPoints; i++) diag[i] = (double)s.colDiag[i];
    for (int i = 0; i < 3 * s.simPoints; i++) rhs[i] = (double)s.colRhs[i];
    return s.cnt.activePairs;
  });
}
// local projection of x with the current initRhs (after a step, or zero) and no collision term
int orc_debug_local(void* hp, const double* x, const double* initRhs, double* rhs) {
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [&](auto& s) {
    using R = typename std::remove_reference<decltype(s.X[0])>::type;
    std::vector<R> a(x, x + 3 * s.totalPoints), out(3 * s.simPoints);
    for (int i = 0; i < 3 * s.simPoints; i++) s.initRhs[i] = initRhs[i];
    bool act = s.activeCollision;
    s.activeCollision = false;
    s.local_projection(a, out);
    s.activeCollision = act;
    for (int i = 0; i < 3 * s.simPoints; i++) rhs[i] = (double)out[i];
    return 0;
  });
}
// one Chebyshev solve: diag (sim), rhs (3*sim), x0 (3*total) -> x (3*sim)
int orc_debug_solve(void* hp, const double* diag, const double* rhs, const double* x0, double* xout) {
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [&](auto& s) {
    using R = typename std::remove_reference<decltype(s.X[0])>::type;
    for (int i = 0; i < s.simPoints; i++) s.sysDiag[i] = diag[i];
    for (int i = 0; i < 3 * s.simPoints; i++) s.sysRhs[i] = rhs[i];
    s.update_operators();
    std::vector<R> A(x0, x0 + 3 * s.totalPoints), B = A, C = A;
    std::vector<R>* pre = &B;
    std::vector<R>* cur = &A;
    std::vector<R>* next = &C;
    s.cheb_solve(pre, cur, next);
    for (int i = 0; i < 3 * s.simPoints; i++) xout[i] = (double)(*cur)[i];
    return 0;
  });
}
int orc_elastic_diag(void* hp, double* out) {
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [&](auto& s) { for (int i = 0; i < s.simPoints; i++) out[i] = (double)s.elasticDiag[i]; return 0; });
}
int orc_mass(void* hp, double* out) {
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [&](auto& s) { for (int i = 0; i < s.simPoints; i++) out[i] = (double)s.mass[i]; return 0; });
}

// sysHess = elastic + M/dt^2 as CSR (rows ascending, columns ascending);
// returns nnz (call with null to size).  Rebuilt from R1 and the diagonal.
int orc_sys_matrix(void* hp, int* rowOff, int* cols, double* vals) {
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [&](auto& s) {
    int nnz = 0;
    const double dt = s.prm.v[P_DT];
    for (int i = 0; i < s.simPoints; i++) {
      if (rowOff) rowOff[i] = nnz;
      // R1 row: slot 0 is the diagonal (value 0), the rest are -offdiag
      std::vector<std::pair<int, double>> row;
      row.push_back({i, (double)s.elasticDiag[i] + (double)s.mass[i] / (dt * dt)});
      for (size_t a = 1; a < s.r1Idx[i].size(); a++) row.push_back({s.r1Idx[i][a], -(double)s.r1Val[i][a]});
      std::sort(row.begin(), row.end());
      for (auto& e : row) {
        if (cols) cols[nnz] = e.first;
        if (vals) vals[nnz] = e.second;
        nnz++;
      }
    }
    if (rowOff) rowOff[s.simPoints] = nnz;
    return nnz;
  });
}
// global boundary faces (3 point ids each) and patches (count, then per patch
// its face count and global face ids); return the int counts
int orc_boundary_faces(void* hp, int* out) {
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [&](auto& s) {
    if (out)
      for (size_t i = 0; i < s.bFaces.size(); i++)
        for (int k = 0; k < 3; k++) out[3 * i + k] = s.bFaces[i][k];
    return (int)(3 * s.bFaces.size());
  });
}
int orc_patches(void* hp, int* out) {
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [&](auto& s) {
    std::vector<int> b;
    b.push_back((int)s.patchFaces.size());
    for (size_t p = 0; p < s.patchFaces.size(); p++) {
      b.push_back((int)s.patchFaces[p].size());
      for (int f : s.patchFaces[p]) b.push_back(f);
    }
    if (out) std::copy(b.begin(), b.end(), out);
    return (int)b.size();
  });
}

// model-local boundary.dat (PdIpcModel.h:219-251) and patchInfo.dat (:291-318)
// images for the simulated model `mi`; returns the byte count (call with null to size)
static int serialize_model(const Model& m, int which, int* out) {
  std::vector<int> b;
  if (which == 0) {
    b.push_back((int)m.bPoints.size());
    for (int p : m.bPoints) b.push_back(p);
    b.push_back((int)m.bEdges.size());
    for (auto& e : m.bEdges) { b.push_back(e[0]); b.push_back(e[1]); }
    b.push_back((int)m.bFaces.size());
    for (auto& f : m.bFaces) { b.push_back(f[0]); b.push_back(f[1]); b.push_back(f[2]); }
  } else {
    b.push_back((int)m.bPoints.size());
    b.push_back((int)m.bEdges.size());
    b.push_back((int)m.bFaces.size());
    b.push_back((int)m.patches.size());
    for (auto& p : m.patches) {
      b.push_back((int)p.size());
      for (int f : p) b.push_back(f);
    }
  }
  if (out) std::copy(b.begin(), b.end(), out);
  return (int)b.size();
}
int orc_model_dat(void* hp, int mi, int which, int* out) {
  auto* h = (OracleHandle*)hp;
  return with_sim(h, [&](auto& s) {
    if (mi < 0 || mi >= (int)s.models.size()) return -1;
    return serialize_model(s.models[mi], which, out);
  });
}

// math entry points for the closed-form / LAPACK / brute-force pinning tests
void orc_math_strain_projection(const double* F, double lo, double hi, double* P) { strain_projection<double>(F, lo, hi, P); }
void orc_math_svd(const double* F, double* U, double* s, double* V) { svd3<double>(F, U, s, V); }
double orc_math_pt_dist2(const double* p, const double* a, const double* b, const double* c, double* bary) {
  double dir[3];
  return vertex_triangle_sqd<double>(p, a, b, c, &bary[0], &bary[1], &bary[2], dir);
}
double orc_math_ee_dist2(const double* a0, const double* a1, const double* b0, const double* b1, double* rs) {
  double dir[3];
  return edge_edge_sqd<double>(a0, a1, b0, b1, &rs[0], &rs[1], dir);
}
double orc_math_ccd(int type, const double* x, const double* dx, int touch_policy, int iters, double step,
                    double gap) {
  double X[4][3], D[4][3];
  for (int i = 0; i < 4; i++)
    for (int k = 0; k < 3; k++) { X[i][k] = x[3 * i + k]; D[i][k] = dx[3 * i + k]; }
  return ccd_toi<double>(type, X, D, iters, step, gap, touch_policy);
}
double orc_math_barrier(double d2, double dh2, double kappa) { return barrier_value<double>(d2, dh2, kappa); }

}  // extern "C"
