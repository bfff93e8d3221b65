// Host precompute for the PD-IPC step (see host_setup.hpp for the reference map).
#include "host_setup.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <sched.h>
#include <stdexcept>
#include <thread>

namespace pdipc {
namespace {

// Host threads for the row-parallel parts of the precompute: OMP_NUM_THREADS
// when set, else the CPUs this process may run on (at most 64).
int setup_threads() {
  const char* e = std::getenv("OMP_NUM_THREADS");
  int n = e ? std::atoi(e) : 0;
  if (n <= 0) {
    cpu_set_t set;
    n = sched_getaffinity(0, sizeof(set), &set) == 0 ? CPU_COUNT(&set) : 1;
  }
  return std::max(1, std::min(n, 64));
}

// fn(begin, end, t) over `parts` contiguous ranges of [0, n), one thread each
template <class F>
void parallel_ranges(int64_t n, int parts, F fn) {
  parts = (int)std::max<int64_t>(1, std::min<int64_t>(parts, n / 1024 + 1));
  std::vector<std::thread> th;
  for (int t = 0; t < parts; t++) {
    int64_t b = n * t / parts, e = n * (t + 1) / parts;
    if (t == parts - 1) fn(b, e, t);
    else th.emplace_back(fn, b, e, t);
  }
  for (auto& x : th) x.join();
}

inline void sub3(const double* a, const double* b, double* r) { r[0] = a[0] - b[0]; r[1] = a[1] - b[1]; r[2] = a[2] - b[2]; }
inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void cross3(const double* a, const double* b, double* r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
inline uint64_t ekey(int a, int b) { return a < b ? ((uint64_t)(uint32_t)a << 32) | (uint32_t)b : ((uint64_t)(uint32_t)b << 32) | (uint32_t)a; }

// CSR helper: counts -> offsets
void counts_to_offsets(std::vector<int>& off) {
  int s = 0;
  for (auto& v : off) { int c = v; v = s; s += c; }
  off.push_back(s);
}

// Local collision topology of one model (PdIpcModel.h:22-37).
struct LocalTopo {
  std::vector<int> bPoints;
  std::vector<int> bEdges;  // 2 per edge, sorted pair
  std::vector<int> bFaces;  // 3 per face
  std::vector<std::vector<int>> patches;
  std::vector<int> faceEdgesOff, faceEdges;    // owned edges per face (local edge ids)
  std::vector<int> facePointsOff, facePoints;  // owned points per face (local point ids)
};

// Sorted unique edge list of a triangle mesh (BaseSurfaceMesh.cpp:219-241).
std::vector<uint64_t> surface_edges(const std::vector<int>& F) {
  std::vector<uint64_t> e;
  e.reserve(F.size());
  for (size_t j = 0; j < F.size() / 3; j++)
    for (int k = 0; k < 3; k++) e.push_back(ekey(F[3 * j + k], F[3 * j + (k + 1) % 3]));
  std::sort(e.begin(), e.end());
  e.erase(std::unique(e.begin(), e.end()), e.end());
  return e;
}

inline int edge_id(const std::vector<uint64_t>& edges, int a, int b) {
  uint64_t k = ekey(a, b);
  auto it = std::lower_bound(edges.begin(), edges.end(), k);
  return (int)(it - edges.begin());
}

// Patch grid of PdIpcClothModel::setCollisionPatches (PdIpcClothModel.h:45-140), in float like the reference.
void cloth_patches(const ModelInput& m, LocalTopo& t) {
  int nv = m.nv();
  float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int i : t.bPoints)
    for (int k = 0; k < 3; k++) {
      float p = (float)m.verts[3 * i + k];
      mn[k] = p < mn[k] ? p : mn[k];
      mx[k] = p > mx[k] ? p : mx[k];
    }
  float sq = std::sqrt((float)nv);
  int len = sq < (float)m.patchDim ? (int)sq : m.patchDim;
  if (len < 1) len = 1;
  float size[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
  int u = 0, v = 2;  // used when no extent is strictly smallest
  if (size[0] < size[1] && size[0] < size[2]) { u = 1; v = 2; }
  if (size[1] < size[0] && size[1] < size[2]) { u = 0; v = 2; }
  if (size[2] < size[0] && size[2] < size[1]) { u = 0; v = 1; }
  const float du = size[u] / len, dv = size[v] / len;
  // the reference scans the cells in (r, c) order and takes the first whose
  // closed box holds the centroid; boxes are [mn + r*du, mn + r*du + du]
  std::vector<float> lou(len), hiu(len), lov(len), hiv(len);
  for (int r = 0; r < len; r++) { lou[r] = mn[u] + r * du; hiu[r] = lou[r] + du; }
  for (int c = 0; c < len; c++) { lov[c] = mn[v] + c * dv; hiv[c] = lov[c] + dv; }
  std::vector<std::vector<int>> pocket((size_t)len * len);
  int nf = (int)t.bFaces.size() / 3;
  for (int i = 0; i < nf; i++) {
    float cent[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++)
      for (int j = 0; j < 3; j++) cent[j] += (float)m.verts[3 * t.bFaces[3 * i + k] + j];
    for (int j = 0; j < 3; j++) cent[j] /= 3;
    // first r (then c) whose closed interval contains the centroid
    for (int r = 0; r < len; r++) {
      if (!(cent[u] >= lou[r] && cent[u] <= hiu[r])) continue;
      int found = -1;
      for (int c = 0; c < len; c++)
        if (cent[v] >= lov[c] && cent[v] <= hiv[c]) { found = c; break; }
      if (found >= 0) { pocket[(size_t)r * len + found].push_back(i); break; }
    }
  }
  for (auto& p : pocket)
    if (!p.empty()) t.patches.push_back(std::move(p));
}

// 3-D patch grid of PdIpcVolumetricModel::setCollisionPatches (PdIpcVolumetricModel.h:115-181), float.
void tet_patches(const ModelInput& m, LocalTopo& t) {
  float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int i : t.bPoints)
    for (int k = 0; k < 3; k++) {
      float p = (float)m.verts[3 * i + k];
      mn[k] = p < mn[k] ? p : mn[k];
      mx[k] = p > mx[k] ? p : mx[k];
    }
  int len = m.patchDim < 1 ? 1 : m.patchDim;
  float size[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
  std::vector<float> lo[3], hi[3];
  for (int k = 0; k < 3; k++) {
    lo[k].resize(len); hi[k].resize(len);
    for (int s = 0; s < len; s++) { lo[k][s] = mn[k] + s * (size[k] / len); hi[k][s] = lo[k][s] + size[k] / len; }
  }
  std::vector<std::vector<int>> pocket((size_t)len * len * len);
  int nf = (int)t.bFaces.size() / 3;
  for (int i = 0; i < nf; i++) {
    float cent[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++)
      for (int j = 0; j < 3; j++) cent[j] += (float)m.verts[3 * t.bFaces[3 * i + k] + j];
    for (int j = 0; j < 3; j++) cent[j] /= 3;
    bool done = false;
    for (int s = 0; s < len && !done; s++) {
      if (!(cent[0] >= lo[0][s] && cent[0] <= hi[0][s])) continue;
      for (int r = 0; r < len && !done; r++) {
        if (!(cent[1] >= lo[1][r] && cent[1] <= hi[1][r])) continue;
        for (int c = 0; c < len; c++)
          if (cent[2] >= lo[2][c] && cent[2] <= hi[2][c]) {
            pocket[((size_t)s * len + r) * len + c].push_back(i);
            done = true;
            break;
          }
      }
    }
  }
  for (auto& p : pocket)
    if (!p.empty()) t.patches.push_back(std::move(p));
}

// Ownership rule of PdIpcModel::postProcessBoundaryInfo (PdIpcModel.h:84-172):
// each boundary edge / point goes to its incident face with the fewest owned
// items so far (first in face order on ties).
void assign_ownership(int nvLocal, LocalTopo& t, const std::vector<uint64_t>& edgeKeys) {
  int nf = (int)t.bFaces.size() / 3, ne = (int)t.bEdges.size() / 2, np = (int)t.bPoints.size();
  std::vector<int> pointSlot(nvLocal, -1);
  for (int i = 0; i < np; i++) pointSlot[t.bPoints[i]] = i;
  // incident faces, in face order
  std::vector<int> pfOff(np, 0), efOff(ne, 0);
  for (int i = 0; i < nf; i++)
    for (int k = 0; k < 3; k++) {
      pfOff[pointSlot[t.bFaces[3 * i + k]]]++;
      efOff[edge_id(edgeKeys, t.bFaces[3 * i + k], t.bFaces[3 * i + (k + 1) % 3])]++;
    }
  counts_to_offsets(pfOff);
  counts_to_offsets(efOff);
  std::vector<int> pf(pfOff.back()), ef(efOff.back()), pfc(np, 0), efc(ne, 0);
  for (int i = 0; i < nf; i++)
    for (int k = 0; k < 3; k++) {
      int p = pointSlot[t.bFaces[3 * i + k]];
      pf[pfOff[p] + pfc[p]++] = i;
      int e = edge_id(edgeKeys, t.bFaces[3 * i + k], t.bFaces[3 * i + (k + 1) % 3]);
      ef[efOff[e] + efc[e]++] = i;
    }
  // The reference lists a face's three edges as (0,1), (1,2), (2,0); the order
  // within one face does not change which face is first, only duplicates matter.
  std::vector<std::vector<int>> fe(nf), fp(nf);
  for (int e = 0; e < ne; e++) {
    int best = ef[efOff[e]];
    size_t bs = fe[best].size();
    for (int k = efOff[e] + 1; k < efOff[e + 1]; k++)
      if (fe[ef[k]].size() < bs) { best = ef[k]; bs = fe[best].size(); }
    fe[best].push_back(e);
  }
  for (int p = 0; p < np; p++) {
    int best = pf[pfOff[p]];
    size_t bs = fp[best].size();
    for (int k = pfOff[p] + 1; k < pfOff[p + 1]; k++)
      if (fp[pf[k]].size() < bs) { best = pf[k]; bs = fp[best].size(); }
    fp[best].push_back(t.bPoints[p]);
  }
  t.faceEdgesOff.assign(1, 0);
  t.facePointsOff.assign(1, 0);
  for (int i = 0; i < nf; i++) {
    t.faceEdges.insert(t.faceEdges.end(), fe[i].begin(), fe[i].end());
    t.faceEdgesOff.push_back((int)t.faceEdges.size());
    t.facePoints.insert(t.facePoints.end(), fp[i].begin(), fp[i].end());
    t.facePointsOff.push_back((int)t.facePoints.size());
  }
}

void surface_topo(const ModelInput& m, LocalTopo& t, std::vector<uint64_t>& edgeKeys) {
  edgeKeys = surface_edges(m.elems);
  int nv = m.nv();
  t.bPoints.resize(nv);
  std::iota(t.bPoints.begin(), t.bPoints.end(), 0);
  t.bEdges.resize(2 * edgeKeys.size());
  for (size_t i = 0; i < edgeKeys.size(); i++) {
    t.bEdges[2 * i] = (int)(edgeKeys[i] >> 32);
    t.bEdges[2 * i + 1] = (int)(edgeKeys[i] & 0xffffffffu);
  }
  t.bFaces = m.elems;
  cloth_patches(m, t);
  assign_ownership(nv, t, edgeKeys);
}

// Boundary of a tet mesh: faces used by exactly one tet, oriented outward,
// listed in ascending order of their sorted vertex triple.
void tet_topo(const ModelInput& m, LocalTopo& t, std::vector<uint64_t>& edgeKeys) {
  static const int eb[4][3] = {{3, 1, 0}, {3, 2, 1}, {3, 0, 2}, {2, 1, 0}};
  int nt = (int)m.elems.size() / 4;
  struct FaceRec { int s[3]; int f[3]; int tet; };
  std::vector<FaceRec> recs;
  recs.reserve(4 * (size_t)nt);
  for (int e = 0; e < nt; e++)
    for (int k = 0; k < 4; k++) {
      FaceRec r;
      for (int j = 0; j < 3; j++) r.f[j] = m.elems[4 * e + eb[k][j]];
      r.s[0] = r.f[0]; r.s[1] = r.f[1]; r.s[2] = r.f[2];
      std::sort(r.s, r.s + 3);
      r.tet = e;
      recs.push_back(r);
    }
  std::sort(recs.begin(), recs.end(), [](const FaceRec& a, const FaceRec& b) {
    if (a.s[0] != b.s[0]) return a.s[0] < b.s[0];
    if (a.s[1] != b.s[1]) return a.s[1] < b.s[1];
    return a.s[2] < b.s[2];
  });
  const double* V = m.verts.data();
  for (size_t i = 0; i < recs.size();) {
    size_t j = i + 1;
    while (j < recs.size() && recs[j].s[0] == recs[i].s[0] && recs[j].s[1] == recs[i].s[1] && recs[j].s[2] == recs[i].s[2]) j++;
    if (j - i == 1) {
      const FaceRec& r = recs[i];
      int face[3] = {r.f[0], r.f[1], r.f[2]};
      double cent[3] = {0, 0, 0}, fc[3] = {0, 0, 0};
      for (int k = 0; k < 3; k++) {
        for (int q = 0; q < 4; q++) cent[k] += V[3 * m.elems[4 * r.tet + q] + k];
        cent[k] /= 4;
        for (int q = 0; q < 3; q++) fc[k] += V[3 * face[q] + k];
        fc[k] /= 3;
      }
      double e1[3], e2[3], n[3], cd[3];
      sub3(V + 3 * face[1], V + 3 * face[0], e1);
      sub3(V + 3 * face[2], V + 3 * face[0], e2);
      cross3(e1, e2, n);
      sub3(cent, fc, cd);
      if (dot3(cd, n) > 0) std::swap(face[1], face[2]);
      t.bFaces.insert(t.bFaces.end(), face, face + 3);
    }
    i = j;
  }
  edgeKeys = surface_edges(t.bFaces);
  std::vector<int> pts(t.bFaces.begin(), t.bFaces.end());
  std::sort(pts.begin(), pts.end());
  pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
  t.bPoints = pts;
  t.bEdges.resize(2 * edgeKeys.size());
  for (size_t i = 0; i < edgeKeys.size(); i++) {
    t.bEdges[2 * i] = (int)(edgeKeys[i] >> 32);
    t.bEdges[2 * i + 1] = (int)(edgeKeys[i] & 0xffffffffu);
  }
  tet_patches(m, t);
  assign_ownership(m.nv(), t, edgeKeys);
}

std::vector<int> dat_boundary(const LocalTopo& t) {
  std::vector<int> b;
  b.push_back((int)t.bPoints.size());
  b.insert(b.end(), t.bPoints.begin(), t.bPoints.end());
  b.push_back((int)t.bEdges.size() / 2);
  b.insert(b.end(), t.bEdges.begin(), t.bEdges.end());
  b.push_back((int)t.bFaces.size() / 3);
  b.insert(b.end(), t.bFaces.begin(), t.bFaces.end());
  return b;
}
std::vector<int> dat_patches(const LocalTopo& t) {
  std::vector<int> b = {(int)t.bPoints.size(), (int)t.bEdges.size() / 2, (int)t.bFaces.size() / 3, (int)t.patches.size()};
  for (auto& p : t.patches) {
    b.push_back((int)p.size());
    b.insert(b.end(), p.begin(), p.end());
  }
  return b;
}

struct Triplet { int r, c; double v; };

}  // namespace

std::string build_host_setup(const std::vector<ModelInput>& sims, const std::vector<ModelInput>& statics,
                             const StepConstants& k, HostSetup& hs) {
  hs = HostSetup();
  hs.clothTriBase.assign(sims.size(), -1);
  const double dt = k.dt;
  std::vector<LocalTopo> topo(sims.size() + statics.size());
  std::vector<int> base(sims.size() + statics.size());
  std::vector<Triplet> trip;
  // ---- simulated models (PdIpcSimulator.cpp:69-124)
  int nPts = 0;
  for (size_t mi = 0; mi < sims.size(); mi++) {
    const ModelInput& m = sims[mi];
    int nv = m.nv();
    if (nv <= 0) return "empty model";
    base[mi] = nPts;
    std::vector<uint64_t> edgeKeys;
    LocalTopo& t = topo[mi];
    if (m.kind == MODEL_CLOTH) {
      if (m.elems.size() % 3 || m.elems.empty()) return "cloth needs triangles";
      for (int v : m.elems) if (v < 0 || v >= nv) return "triangle index out of range";
      surface_topo(m, t, edgeKeys);
      int nf = (int)m.elems.size() / 3, ne = (int)edgeKeys.size();
      const double* V = m.verts.data();
      // lumped mass (PdIpcClothModel.h:177-195)
      std::vector<double> area(nv, 0.0);
      for (int f = 0; f < nf; f++) {
        const int* F = &m.elems[3 * f];
        double e1[3], e2[3], n[3];
        sub3(V + 3 * F[1], V + 3 * F[0], e1);
        sub3(V + 3 * F[2], V + 3 * F[0], e2);
        cross3(e1, e2, n);
        double a = 0.5 * std::fabs(std::sqrt(dot3(n, n)));
        for (int q = 0; q < 3; q++) area[F[q]] += a / 3;
      }
      for (int v = 0; v < nv; v++) {
        if (!(area[v] > 0)) return "cloth vertex with zero area (unreferenced vertex)";
        double ms = area[v] * m.density;
        hs.mass.push_back(ms);
        hs.invMass.push_back(1.0 / ms);
      }
      // cotangent weights per edge (PdIpcClothModel.h:220-262); faces of an edge in face order
      std::vector<int> efCount(ne, 0);
      std::vector<std::array<int, 2>> efList(ne, {-1, -1});
      for (int f = 0; f < nf; f++)
        for (int q = 0; q < 3; q++) {
          int e = edge_id(edgeKeys, m.elems[3 * f + q], m.elems[3 * f + (q + 1) % 3]);
          if (efCount[e] < 2) efList[e][efCount[e]] = f;
          efCount[e]++;
        }
      // spring weight per edge: s0, or the mean stiffness of the adjacent triangles
      std::vector<double> springW(ne, m.s0);
      int triBase = hs.nClothTris;
      int edgeBase = (int)hs.edgeTriOff.size();
      if (!m.triK.empty() && (int)m.triK.size() != nf) return "per-triangle stiffness does not match the cloth";
      for (int e = 0; e < ne; e++) {
        int c = std::min(efCount[e], 2);
        if (!m.triK.empty()) {
          double acc = 0;
          for (int q = 0; q < c; q++) acc += m.triK[efList[e][q]];
          springW[e] = acc / c;
        }
        hs.edgeTriOff.push_back((int)hs.edgeTriList.size());
        for (int q = 0; q < c; q++) hs.edgeTriList.push_back(triBase + efList[e][q]);
      }
      hs.clothTriBase[mi] = triBase;
      hs.nClothTris += nf;
      std::vector<double> cotw(ne, 0.0);
      for (int e = 0; e < ne; e++) {
        int a = (int)(edgeKeys[e] >> 32), b = (int)(edgeKeys[e] & 0xffffffffu);
        if (efCount[e] > 2) return "non-manifold cloth edge";
        for (int s = 0; s < efCount[e]; s++) {
          const int* F = &m.elems[3 * efList[e][s]];
          int opp = -1;
          for (int q = 0; q < 3; q++)
            if (F[q] != a && F[q] != b) { opp = F[q]; break; }
          double x10[3], x20[3];
          sub3(V + 3 * a, V + 3 * opp, x10);
          sub3(V + 3 * b, V + 3 * opp, x20);
          double l10 = std::sqrt(dot3(x10, x10)), l20 = std::sqrt(dot3(x20, x20));
          double th = std::acos(dot3(x10, x20) / (l10 * l20));
          cotw[e] += -0.5 * (1.0 / std::tan(th));
        }
      }
      // per-vertex ring: neighbours in edge order = ascending vertex id (PdIpcClothModel.h:271-318)
      std::vector<int> ringOff(nv + 1, 0);
      for (int e = 0; e < ne; e++) { ringOff[edgeKeys[e] >> 32]++; ringOff[edgeKeys[e] & 0xffffffffu]++; }
      {
        int s = 0;
        for (int v = 0; v < nv; v++) { int c = ringOff[v]; ringOff[v] = s; s += c; }
        ringOff[nv] = s;
      }
      std::vector<int> ringE(ringOff[nv]), fill(nv, 0);
      for (int e = 0; e < ne; e++) {
        int a = (int)(edgeKeys[e] >> 32), b = (int)(edgeKeys[e] & 0xffffffffu);
        ringE[ringOff[a] + fill[a]++] = e;
        ringE[ringOff[b] + fill[b]++] = e;
      }
      int cbase = (int)hs.cCenter.size();
      std::vector<std::vector<double>> lboRows(nv);
      for (int v = 0; v < nv; v++) {
        int nb = ringOff[v + 1] - ringOff[v];
        hs.cCenter.push_back(base[mi] + v);
        if (hs.cOff.empty()) hs.cOff.push_back(0);
        std::vector<double> lbo(nb + 1, 0.0);
        double sum = 0, curv[3] = {0, 0, 0};
        hs.cNbr.push_back(base[mi] + v);
        hs.cRest.push_back(0.0);
        hs.cStrain.push_back(m.s0);
        hs.cSlotEdge.push_back(-1);
        for (int q = 0; q < nb; q++) {
          int e = ringE[ringOff[v] + q];
          int a = (int)(edgeKeys[e] >> 32), b = (int)(edgeKeys[e] & 0xffffffffu);
          int ve = v == a ? b : a;
          double d[3];
          sub3(V + 3 * v, V + 3 * ve, d);
          hs.cNbr.push_back(base[mi] + ve);
          hs.cRest.push_back(std::sqrt(dot3(d, d)));
          hs.cStrain.push_back(springW[e]);
          hs.cSlotEdge.push_back(edgeBase + e);
          double ew = cotw[e] / area[v];
          sum += ew;
          lbo[q + 1] = ew;
        }
        lbo[0] = -sum;
        for (int q = 0; q <= nb; q++) {
          int pid = hs.cNbr[hs.cOff.back() + q] - base[mi];
          for (int j = 0; j < 3; j++) curv[j] += lbo[q] * V[3 * pid + j];
        }
        hs.cLbo.insert(hs.cLbo.end(), lbo.begin(), lbo.end());
        for (int q = 0; q <= nb; q++) hs.cBendOut.push_back(lbo[q] * m.s1 * area[v]);
        hs.cH0.push_back(std::sqrt(dot3(curv, curv)));
        hs.cOff.push_back((int)hs.cNbr.size());
        lboRows[v] = lbo;
        // PD system: springs (PdIpcSimulator.cpp:358-369) and bending area*bend * l l^T (:371-396)
        int p0 = base[mi] + v;
        for (int q = 0; q < nb; q++) {
          double w = springW[ringE[ringOff[v] + q]];
          trip.push_back({p0, p0, w});
          trip.push_back({p0, hs.cNbr[hs.cOff[hs.cOff.size() - 2] + q + 1], -w});
        }
        double ba = area[v] * m.s1;
        int o = hs.cOff[hs.cOff.size() - 2];
        for (int r = 0; r <= nb; r++)
          for (int c = 0; c <= nb; c++) trip.push_back({hs.cNbr[o + r], hs.cNbr[o + c], ba * (lbo[r] * lbo[c])});
      }
      // bending scatter as a gather: slot s of vertex i receives bend(k -> i) from k = ring(i)[s]
      // where bend(k -> i) = L_k,i * bend * area_k (PdIpcClothModel.h:315)
      for (int v = 0; v < nv; v++) {
        int o = hs.cOff[cbase + v];
        int nb = ringOff[v + 1] - ringOff[v];
        hs.cBendIn.push_back(lboRows[v][0] * m.s1 * area[v]);
        hs.cLboIn.push_back(lboRows[v][0]);
        for (int q = 0; q < nb; q++) {
          int kv = hs.cNbr[o + q + 1] - base[mi];
          // slot of v in k's ring
          int ko = ringOff[kv], kn = ringOff[kv + 1] - ringOff[kv], slot = -1;
          for (int z = 0; z < kn; z++) {
            int e = ringE[ko + z];
            int a = (int)(edgeKeys[e] >> 32), b = (int)(edgeKeys[e] & 0xffffffffu);
            if ((a == v && b == kv) || (b == v && a == kv)) { slot = z + 1; break; }
          }
          hs.cBendIn.push_back(lboRows[kv][slot] * m.s1 * area[kv]);
          hs.cLboIn.push_back(lboRows[kv][slot]);
        }
      }
      for (int f : m.fixed) {
        if (f < 0 || f >= nv) return "fixed id out of range";
      }
      hs.nCloth += nv;
    } else if (m.kind == MODEL_TET) {
      if (m.elems.size() % 4 || m.elems.empty()) return "tet model needs tets";
      for (int v : m.elems) if (v < 0 || v >= nv) return "tet index out of range";
      tet_topo(m, t, edgeKeys);
      int nt = (int)m.elems.size() / 4;
      const double* V = m.verts.data();
      std::vector<double> vol(nv, 0.0);
      for (int e = 0; e < nt; e++) {
        const int* T = &m.elems[4 * e];
        double E[3][3];
        for (int r = 0; r < 3; r++) sub3(V + 3 * T[r + 1], V + 3 * T[0], E[r]);
        double det = E[0][0] * (E[1][1] * E[2][2] - E[1][2] * E[2][1]) - E[0][1] * (E[1][0] * E[2][2] - E[1][2] * E[2][0]) +
                     E[0][2] * (E[1][0] * E[2][1] - E[1][1] * E[2][0]);
        if (!(std::fabs(det) > 0)) return "degenerate tet";
        double tv = std::fabs(det) / 6.0;
        for (int q = 0; q < 4; q++) vol[T[q]] += tv;
        // Dm^{-1} row-major = column-major data of (Dm^T)^{-1} (PdIpcSimulator.cpp:318-329)
        double inv[3][3];
        inv[0][0] = (E[1][1] * E[2][2] - E[1][2] * E[2][1]) / det;
        inv[0][1] = (E[0][2] * E[2][1] - E[0][1] * E[2][2]) / det;
        inv[0][2] = (E[0][1] * E[1][2] - E[0][2] * E[1][1]) / det;
        inv[1][0] = (E[1][2] * E[2][0] - E[1][0] * E[2][2]) / det;
        inv[1][1] = (E[0][0] * E[2][2] - E[0][2] * E[2][0]) / det;
        inv[1][2] = (E[0][2] * E[1][0] - E[0][0] * E[1][2]) / det;
        inv[2][0] = (E[1][0] * E[2][1] - E[1][1] * E[2][0]) / det;
        inv[2][1] = (E[0][1] * E[2][0] - E[0][0] * E[2][1]) / det;
        inv[2][2] = (E[0][0] * E[1][1] - E[0][1] * E[1][0]) / det;
        double idm[9];
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++) idm[3 * r + c] = inv[c][r];
        double w = m.s0 * tv;
        for (int q = 0; q < 4; q++) hs.tetIdx.push_back(base[mi] + T[q]);
        hs.tetDminv.insert(hs.tetDminv.end(), idm, idm + 9);
        hs.tetW.push_back(w);
        // G^T Dm^{-1} Dm^{-T} G (PdIpcSimulator.cpp:331-350)
        double half[3][4];
        for (int r = 0; r < 3; r++) {
          half[r][0] = -idm[r] - idm[3 + r] - idm[6 + r];
          half[r][1] = idm[r];
          half[r][2] = idm[3 + r];
          half[r][3] = idm[6 + r];
        }
        for (int r = 0; r < 4; r++)
          for (int c = 0; c < 4; c++)
            trip.push_back({base[mi] + T[r], base[mi] + T[c], (half[0][r] * half[0][c] + half[1][r] * half[1][c] + half[2][r] * half[2][c]) * w});
      }
      for (int v = 0; v < nv; v++) {
        if (!(vol[v] > 0)) return "tet-mesh vertex in no tet";
        hs.mass.push_back(vol[v] * m.density);
        hs.invMass.push_back(1.0 / (vol[v] * m.density));
      }
      for (int f : m.fixed) if (f < 0 || f >= nv) return "fixed id out of range";
    } else {
      return "simulated model of unknown kind";
    }
    hs.boundaryDat.push_back(dat_boundary(t));
    hs.patchDat.push_back(dat_patches(t));
    nPts += nv;
  }
  hs.nSim = nPts;
  hs.nTets = (int)hs.tetW.size();
  for (size_t si = 0; si < statics.size(); si++) {
    const ModelInput& m = statics[si];
    size_t mi = sims.size() + si;
    base[mi] = nPts;
    if (m.elems.size() % 3 || m.elems.empty()) return "static model needs triangles";
    for (int v : m.elems) if (v < 0 || v >= m.nv()) return "triangle index out of range";
    std::vector<uint64_t> edgeKeys;
    surface_topo(m, topo[mi], edgeKeys);
    nPts += m.nv();
  }
  hs.nTotal = nPts;
  // positions (PdIpcSimulator.cpp:182-194)
  hs.X.resize(3 * (size_t)hs.nTotal);
  {
    size_t o = 0;
    for (auto& m : sims) { std::copy(m.verts.begin(), m.verts.end(), hs.X.begin() + o); o += m.verts.size(); }
    for (auto& m : statics) { std::copy(m.verts.begin(), m.verts.end(), hs.X.begin() + o); o += m.verts.size(); }
  }
  hs.gravY.resize(hs.nSim);
  for (int i = 0; i < hs.nSim; i++) hs.gravY[i] = k.gravity * hs.mass[i];
  hs.clothOfPoint.assign(hs.nSim, -1);
  for (int c = 0; c < (int)hs.cCenter.size(); c++) hs.clothOfPoint[hs.cCenter[c]] = c;
  // position constraints (PdIpcSimulator.cpp:274-299)
  hs.fixK.assign(hs.nSim, 0.0);
  hs.fixFlag.assign(hs.nSim, 0);
  for (size_t mi = 0; mi < sims.size(); mi++)
    for (int f : sims[mi].fixed) {
      int p = base[mi] + f;
      hs.fixFlag[p] = 1;
      hs.fixK[p] = sims[mi].s2;
      trip.push_back({p, p, sims[mi].s2});
    }
  // vertex -> (tet, corner)
  hs.vtOff.assign(hs.nSim, 0);
  for (int e = 0; e < hs.nTets; e++)
    for (int q = 0; q < 4; q++) hs.vtOff[hs.tetIdx[4 * e + q]]++;
  counts_to_offsets(hs.vtOff);
  {
    std::vector<int> fill(hs.nSim, 0);
    hs.vtList.resize(hs.vtOff.back());
    for (int e = 0; e < hs.nTets; e++)
      for (int q = 0; q < 4; q++) {
        int p = hs.tetIdx[4 * e + q];
        hs.vtList[hs.vtOff[p] + fill[p]++] = 4 * e + q;
      }
  }
  // ---- sysHess = elastic + M/dt^2 as sorted, summed triplets
  std::sort(trip.begin(), trip.end(), [](const Triplet& a, const Triplet& b) { return a.r != b.r ? a.r < b.r : a.c < b.c; });
  std::vector<int> hOff(hs.nSim + 1, 0), hCol;
  std::vector<double> hVal;
  for (size_t i = 0; i < trip.size();) {
    size_t j = i;
    double s = 0;
    while (j < trip.size() && trip[j].r == trip[i].r && trip[j].c == trip[i].c) s += trip[j++].v;
    hCol.push_back(trip[i].c);
    hVal.push_back(s);
    hOff[trip[i].r + 1]++;
    i = j;
  }
  trip.clear();
  trip.shrink_to_fit();
  for (int i = 0; i < hs.nSim; i++) hOff[i + 1] += hOff[i];
  hs.elasticDiag.assign(hs.nSim, 0.0);
  for (int i = 0; i < hs.nSim; i++)
    for (int q = hOff[i]; q < hOff[i + 1]; q++)
      if (hCol[q] == i) hs.elasticDiag[i] = hVal[q];
  // ---- R1 = -offdiag(sysHess) with |v| >= 1e-12 (PdIpcSimulator.cpp:640-669)
  const double IG = 1e-12;
  hs.r1Off.assign(1, 0);
  for (int i = 0; i < hs.nSim; i++) {
    for (int q = hOff[i]; q < hOff[i + 1]; q++) {
      int c = hCol[q];
      double v = hVal[q];
      if (c == i) continue;  // M/dt^2 only touches the diagonal
      if (std::fabs(v) < IG) continue;
      hs.r1Col.push_back(c);
      hs.r1Val.push_back(-v);
    }
    hs.r1Off.push_back((int)hs.r1Col.size());
  }
  hs.nnzR1 = (int64_t)hs.r1Col.size() + hs.nSim;
  hs.cSlotR1.assign(hs.cNbr.size(), -1);
  for (size_t c = 0; c < hs.cCenter.size(); c++) {
    int i = hs.cCenter[c];
    for (int sl = hs.cOff[c] + 1; sl < hs.cOff[c + 1]; sl++)
      for (int q = hs.r1Off[i]; q < hs.r1Off[i + 1]; q++)
        if (hs.r1Col[q] == hs.cNbr[sl]) { hs.cSlotR1[sl] = q; break; }
  }
  // ---- R2 = R1 * R1 value pattern, then the 2-path lists (PdIpcSimulator.cpp:704-796).
  // Rows are independent.  Two passes over contiguous row ranges on their own
  // threads: count each row's entries and 2-paths, then (after one allocation
  // of the exact sizes) write them at their offsets; the result is the serial one.
  {
    const int n = hs.nSim;
    const int nt = setup_threads();
    std::vector<int> rowEnt(n, 0);
    std::vector<int64_t> rowPath(n, 0);
    // visit(i, emit): the row's entries in ascending column order; emit(j, r1q, paths)
    // receives the column, the R1 index of (i, j) and the row's valid 2-paths (k, a, ia, ib)
    struct Scratch {
      std::vector<double> acc;
      std::vector<char> seen;
      std::vector<int> touched;
    };
    auto visit = [&](int i, Scratch& S, auto&& onEntry, auto&& onPath) {
      S.touched.clear();
      // the diagonal slot of R1 holds an explicit zero; it only adds structure
      auto touch = [&](int j) { if (!S.seen[j]) { S.seen[j] = 1; S.touched.push_back(j); S.acc[j] = 0.0; } };
      for (int q = hs.r1Off[i]; q < hs.r1Off[i + 1]; q++) touch(hs.r1Col[q]);
      touch(i);
      for (int a = hs.r1Off[i]; a < hs.r1Off[i + 1]; a++) {
        int kk = hs.r1Col[a];
        double vik = hs.r1Val[a];
        touch(kk);
        for (int b = hs.r1Off[kk]; b < hs.r1Off[kk + 1]; b++) {
          int j = hs.r1Col[b];
          touch(j);
          S.acc[j] += vik * hs.r1Val[b];
        }
      }
      std::sort(S.touched.begin(), S.touched.end());
      for (int j : S.touched) {
        S.seen[j] = 0;
        if (std::fabs(S.acc[j]) < IG) continue;
        // 2-paths: k in offdiag(R1 row i) and offdiag(R1 row j), ascending k
        int np = 0;
        int a = hs.r1Off[i], ae = hs.r1Off[i + 1], b = hs.r1Off[j], be = hs.r1Off[j + 1];
        while (a < ae && b < be) {
          if (hs.r1Col[a] < hs.r1Col[b]) a++;
          else if (hs.r1Col[a] > hs.r1Col[b]) b++;
          else {
            double ajk = hs.r1Val[a] * hs.r1Val[b];
            if (!(std::fabs(ajk) < IG)) {
              onPath(hs.r1Col[a], ajk, a, b);
              np++;
            }
            a++; b++;
          }
        }
        if (np == 0) continue;  // no valid 2-path: entry dropped
        int r1q = -1;
        for (int q = hs.r1Off[i]; q < hs.r1Off[i + 1]; q++)
          if (hs.r1Col[q] == j) { r1q = q; break; }
        onEntry(j, r1q, np);
      }
    };
    parallel_ranges(n, nt, [&](int64_t r0, int64_t r1, int) {
      Scratch S{std::vector<double>(n, 0.0), std::vector<char>(n, 0), {}};
      for (int i = (int)r0; i < (int)r1; i++) {
        int ne = 0;
        int64_t np = 0;
        visit(i, S, [&](int, int, int p) { ne++; np += p; }, [&](int, double, int, int) {});
        rowEnt[i] = ne;
        rowPath[i] = np;
      }
    });
    hs.m2Off.assign(n + 1, 0);
    std::vector<int64_t> pathOff(n + 1, 0);
    int64_t nE = 0;
    for (int i = 0; i < n; i++) {
      nE += rowEnt[i];
      if (nE > INT32_MAX) return "the M^2 pattern exceeds 2^31 entries";
      hs.m2Off[i + 1] = (int)nE;
      pathOff[i + 1] = pathOff[i] + rowPath[i];
    }
    int64_t nP = pathOff[n];
    hs.m2Col.resize(nE); hs.m2R1.resize(nE); hs.m2R1Idx.resize(nE); hs.p2Off.resize(nE + 1);
    hs.p2K.resize(nP); hs.p2A.resize(nP); hs.p2Ia.resize(nP); hs.p2Ib.resize(nP);
    hs.p2Off[0] = 0;
    parallel_ranges(n, nt, [&](int64_t r0, int64_t r1, int) {
      Scratch S{std::vector<double>(n, 0.0), std::vector<char>(n, 0), {}};
      for (int i = (int)r0; i < (int)r1; i++) {
        int64_t e = hs.m2Off[i], q = pathOff[i];
        visit(
            i, S,
            [&](int j, int r1q, int) {
              hs.m2Col[e] = j;
              hs.m2R1[e] = r1q >= 0 ? hs.r1Val[r1q] : 0.0;
              hs.m2R1Idx[e] = r1q;
              hs.p2Off[e + 1] = q;
              e++;
            },
            [&](int k, double a, int ia, int ib) {
              hs.p2K[q] = k;
              hs.p2A[q] = a;
              hs.p2Ia[q] = ia;
              hs.p2Ib[q] = ib;
              q++;
            });
      }
    });
    hs.nnzR2 = nE;
    hs.n2paths = nP;
  }
  // ---- collision topology (PdIpcSimulator.cpp:1151-1247)
  {
    int pointOff = 0, faceOff = 0, edgeOff = 0;
    std::vector<int> fEdgeOff(1, 0), fEdge, fPointOff(1, 0), fPoint;
    std::vector<std::vector<int>> patches;
    for (size_t mi = 0; mi < topo.size(); mi++) {
      const LocalTopo& t = topo[mi];
      bool isStatic = mi >= sims.size();
      for (int v : t.bFaces) hs.bFaces.push_back(v + pointOff);
      for (int v : t.bEdges) hs.bEdges.push_back(v + pointOff);
      for (auto& p : t.patches) {
        std::vector<int> q;
        for (int f : p) q.push_back(f + faceOff);
        patches.push_back(q);
        hs.patchStatic.push_back(isStatic ? 1 : 0);
      }
      int nf = (int)t.bFaces.size() / 3;
      for (int f = 0; f < nf; f++) {
        for (int q = t.faceEdgesOff[f]; q < t.faceEdgesOff[f + 1]; q++) fEdge.push_back(t.faceEdges[q] + edgeOff);
        fEdgeOff.push_back((int)fEdge.size());
        for (int q = t.facePointsOff[f]; q < t.facePointsOff[f + 1]; q++) fPoint.push_back(t.facePoints[q] + pointOff);
        fPointOff.push_back((int)fPoint.size());
      }
      pointOff += mi < sims.size() ? sims[mi].nv() : statics[mi - sims.size()].nv();
      faceOff += nf;
      edgeOff += (int)t.bEdges.size() / 2;
    }
    hs.pfOff.assign(1, 0); hs.ppOff.assign(1, 0); hs.peOff.assign(1, 0);
    std::vector<int> es, ps;
    for (auto& p : patches) {
      es.clear(); ps.clear();
      for (int f : p) {
        hs.pfList.push_back(f);
        for (int q = fEdgeOff[f]; q < fEdgeOff[f + 1]; q++) es.push_back(fEdge[q]);
        for (int q = fPointOff[f]; q < fPointOff[f + 1]; q++) ps.push_back(fPoint[q]);
      }
      std::sort(es.begin(), es.end()); es.erase(std::unique(es.begin(), es.end()), es.end());
      std::sort(ps.begin(), ps.end()); ps.erase(std::unique(ps.begin(), ps.end()), ps.end());
      hs.peList.insert(hs.peList.end(), es.begin(), es.end());
      hs.ppList.insert(hs.ppList.end(), ps.begin(), ps.end());
      hs.pfOff.push_back((int)hs.pfList.size());
      hs.peOff.push_back((int)hs.peList.size());
      hs.ppOff.push_back((int)hs.ppList.size());
    }
  }
  hs.colMass.assign(hs.nTotal, 0.0);
  for (int i = 0; i < hs.nSim; i++) hs.colMass[i] = hs.mass[i];
  (void)dt;
  return std::string();
}

// The column union of each tile of TS slices, at most maxCols columns
// spanning at most 65536; per entry its index in the union (Loc: uint16_t or
// uint8_t).  Returns false when a tile does not fit.
template <class Loc>
static bool stage_tiles(const SellLayout& L, int nt, int TS, int maxCols, std::vector<int>& tileOff,
                        std::vector<int>& tileBase, std::vector<uint16_t>& tileCol, std::vector<Loc>& loc,
                        int& stageMax) {
  const int nTiles = (L.nSlices + TS - 1) / TS;
  std::vector<std::vector<int>> uni(nTiles);
  std::vector<uint8_t> ok(nTiles, 1);
  const int64_t total = L.sliceOff.back();
  loc.assign(total, 0);
  parallel_ranges(nTiles, nt, [&](int64_t t0, int64_t t1, int) {
    for (int t = (int)t0; t < (int)t1; t++) {
      const int64_t e0 = L.sliceOff[(size_t)t * TS];
      const int64_t e1 = L.sliceOff[std::min(L.nSlices, (t + 1) * TS)];
      std::vector<int>& U = uni[t];
      U.assign(L.col.begin() + e0, L.col.begin() + e1);
      std::sort(U.begin(), U.end());
      U.erase(std::unique(U.begin(), U.end()), U.end());
      if (U.empty() || U.back() - U.front() > 65535 || (int)U.size() > maxCols) {
        ok[t] = 0;
        continue;
      }
      for (int64_t e = e0; e < e1; e++)
        loc[e] = (Loc)(std::lower_bound(U.begin(), U.end(), L.col[e]) - U.begin());
    }
  });
  if (nTiles == 0 || !std::all_of(ok.begin(), ok.end(), [](uint8_t v) { return v != 0; })) {
    loc.clear();
    return false;
  }
  tileOff.assign(nTiles + 1, 0);
  tileBase.assign(nTiles, 0);
  stageMax = 0;
  for (int t = 0; t < nTiles; t++) {
    tileOff[t + 1] = tileOff[t] + (int)uni[t].size();
    tileBase[t] = uni[t].front();
    stageMax = std::max(stageMax, (int)uni[t].size());
  }
  tileCol.resize(tileOff.back());
  parallel_ranges(nTiles, nt, [&](int64_t t0, int64_t t1, int) {
    for (int t = (int)t0; t < (int)t1; t++)
      for (size_t i = 0; i < uni[t].size(); i++) tileCol[tileOff[t] + i] = (uint16_t)(uni[t][i] - tileBase[t]);
  });
  return true;
}

// The staging tables (see SellLayout): tiles of kTileSlices slices, and for
// lane-major layouts also the narrow tiles of kNarrowTileSlices slices.
static void build_stage_tiles(SellLayout& L, int nt) {
  if (L.sliceOff.back() >= ((int64_t)1 << 29)) return;  // the staged sweeps address entries in 32 bits
  if (L.lanes < 2) return;  // one lane per row: a tile spans 256 rows, not staged
  L.stageOk = stage_tiles(L, nt, kTileSlices, kStageMaxCols, L.tileOff, L.tileBase, L.tileCol, L.loc, L.stageMax);
  if (L.stageOk && L.laneMajor() && L.lanes == 8)
    L.narrowOk = stage_tiles(L, nt, kNarrowTileSlices, kNarrowMaxCols, L.ntileOff, L.ntileBase, L.ntileCol, L.loc8,
                             L.narrowMax);
}

void build_sell(const HostSetup& hs, int lanes, SellLayout& L) {
  L = SellLayout();
  L.lanes = lanes;
  L.rowsPerSlice = 64 / lanes;
  int n = hs.nSim;
  L.nSlices = (n + L.rowsPerSlice - 1) / L.rowsPerSlice;
  const int S = L.rowsPerSlice, nSlots = L.nSlices * S;
  std::vector<int> rowW(n);
  for (int r = 0; r < n; r++) rowW[r] = (hs.m2Off[r + 1] - hs.m2Off[r] + lanes - 1) / lanes;
  auto slice_widths = [&](auto rowOfSlot) {
    L.sliceW.assign(L.nSlices, 0);
    for (int p = 0; p < nSlots; p++) {
      const int r = rowOfSlot(p);
      if (r < n) L.sliceW[p / S] = std::max(L.sliceW[p / S], rowW[r]);
    }
  };
  slice_widths([](int p) { return p; });
  // When padding every slice to the widest costs at most 2% more slots (a
  // regular mesh: only boundary slices are narrower), do so: the sweeps then
  // find a slice's entries by arithmetic instead of a dependent table load.
  int64_t slots = 0, wmax = 0;
  for (int w : L.sliceW) {
    slots += w;
    wmax = std::max<int64_t>(wmax, w);
  }
  if (L.nSlices > 0 && (wmax * L.nSlices - slots) * 50 <= slots) {
    L.uniformW = (int)wmax;
    std::fill(L.sliceW.begin(), L.sliceW.end(), L.uniformW);
  } else if (L.nSlices > 0) {
    // Otherwise order the rows by width (widest first, stable) within windows
    // of kSellSortSlices slices, so that each slice holds rows of similar
    // width (a tet mesh of 5-tet cells alternates two vertex valences from one
    // row to the next: two slices hold one kind each) while the rows of a
    // slice stay near each other; every row keeps its entries and their order,
    // so its sums are unchanged.
    const int win = kSellSortSlices * S;
    L.perm.resize(nSlots);
    for (int p = 0; p < nSlots; p++) L.perm[p] = p;
    for (int w0 = 0; w0 < n; w0 += win)
      std::stable_sort(L.perm.begin() + w0, L.perm.begin() + std::min(n, w0 + win),
                       [&](int a, int b) { return rowW[a] > rowW[b]; });
    L.slotOf.assign(n, 0);
    for (int p = 0; p < n; p++) L.slotOf[L.perm[p]] = p;
    slice_widths([&](int p) { return L.perm[p]; });
  }
  auto rowOf = [&](int p) { return L.perm.empty() ? p : L.perm[p]; };
  L.sliceOff.assign(1, 0);
  for (int s = 0; s < L.nSlices; s++) L.sliceOff.push_back(L.sliceOff.back() + (int64_t)L.sliceW[s] * 64);
  int64_t total = L.sliceOff.back();
  L.col.assign(total, 0);
  L.rowFlag.assign(total, 0);
  L.r1.assign(total, 0.0);
  L.r1Idx.assign(total, -1);
  std::vector<int64_t> src(total, -1);  // CSR entry feeding each SELL slot
  const int nt = setup_threads();
  parallel_ranges(L.nSlices, nt, [&](int64_t s0, int64_t s1, int) {
    for (int s = (int)s0; s < (int)s1; s++)
      for (int lr = 0; lr < L.rowsPerSlice; lr++) {
        int r = rowOf(s * L.rowsPerSlice + lr);
        int rr = r < n ? r : n - 1;  // rows past the end pad with a valid index
        for (int c = 0; c < L.sliceW[s] * lanes; c++) {
          const int64_t e = L.entry(s, lr * lanes + c % lanes, c / lanes);
          L.col[e] = rr;
          L.rowFlag[e] = rr;
          if (r < n && c < hs.m2Off[r + 1] - hs.m2Off[r]) {
            int64_t q = hs.m2Off[r] + c;
            L.col[e] = hs.m2Col[q];
            L.rowFlag[e] = r | (hs.m2Col[q] == r ? (int)0x80000000 : 0);
            L.r1[e] = hs.m2R1[q];
            L.r1Idx[e] = hs.m2R1Idx[q];
            src[e] = q;
          }
        }
      }
  });
  // 16-bit columns relative to the entry's slot (its row, unless rows are
  // reordered: then the sweeps' gathers need not wait for the row index)
  // when every offset fits
  L.col16ok = true;
  for (int s = 0; s < L.nSlices && L.col16ok; s++)
    for (int64_t e = L.sliceOff[s]; e < L.sliceOff[s + 1]; e++) {
      int slot = s * L.rowsPerSlice + L.lane_of(s, e) / lanes;
      int d = L.col[e] - slot;
      if (d < -32768 || d > 32767) {
        L.col16ok = false;
        break;
      }
    }
  if (L.col16ok) {
    L.col16.resize(total);
    parallel_ranges(L.nSlices, nt, [&](int64_t s0, int64_t s1, int) {
      for (int s = (int)s0; s < (int)s1; s++)
        for (int64_t e = L.sliceOff[s]; e < L.sliceOff[s + 1]; e++)
          L.col16[e] = (int16_t)(L.col[e] - (s * L.rowsPerSlice + L.lane_of(s, e) / lanes));
    });
  }
  build_stage_tiles(L, nt);
  L.pOff.assign(total + 1, 0);
  for (int64_t e = 0; e < total; e++) {
    int64_t q = src[e];
    L.pOff[e + 1] = L.pOff[e] + (q >= 0 ? hs.p2Off[q + 1] - hs.p2Off[q] : 0);
  }
  L.pK.resize(L.pOff.back());
  L.pA.resize(L.pOff.back());
  L.pIa.resize(L.pOff.back());
  L.pIb.resize(L.pOff.back());
  parallel_ranges(total, nt, [&](int64_t e0, int64_t e1, int) {
    for (int64_t e = e0; e < e1; e++) {
      int64_t q = src[e];
      if (q < 0) continue;
      std::copy(hs.p2K.begin() + hs.p2Off[q], hs.p2K.begin() + hs.p2Off[q + 1], L.pK.begin() + L.pOff[e]);
      std::copy(hs.p2A.begin() + hs.p2Off[q], hs.p2A.begin() + hs.p2Off[q + 1], L.pA.begin() + L.pOff[e]);
      std::copy(hs.p2Ia.begin() + hs.p2Off[q], hs.p2Ia.begin() + hs.p2Off[q + 1], L.pIa.begin() + L.pOff[e]);
      std::copy(hs.p2Ib.begin() + hs.p2Off[q], hs.p2Ib.begin() + hs.p2Off[q + 1], L.pIb.begin() + L.pOff[e]);
    }
  });
}

}  // namespace pdipc
