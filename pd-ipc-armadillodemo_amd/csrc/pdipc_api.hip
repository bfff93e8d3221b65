// C ABI of the MI355X PD-IPC step (include/pdipc.h) and the step driver
// PdIpcSimulator::advance (Simulator/PD-IPC/PdIpcSimulator.cpp:1683-1796) on one
// HIP stream.  The reference synchronises after every kernel and reads four
// scalars per outer iteration; here kernels are queued back to back and the
// host waits once per outer iteration (for the |d|^2 break test), plus twice in
// the per-step broad phase (to size the pair and record lists).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "handle.hpp"

using namespace pdipc;

namespace pdipc {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

}  // namespace pdipc

namespace pdipc {

// AoS xyz (npoints) -> SoA [3][ld] device
int put_soa(pdipc_handle* h, double* dst, const double* aos, int npts) {
  h->stage.assign(3 * (size_t)h->m.ld, 0.0);
  for (int i = 0; i < npts; i++)
    for (int k = 0; k < 3; k++) h->stage[(size_t)k * h->m.ld + i] = aos[3 * (size_t)i + k];
  HIPCHK(hipMemcpyAsync(dst, h->stage.data(), 3 * (size_t)h->m.ld * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return PDIPC_OK;
}
// partial overwrite: only the first npts points of each component
int put_soa_prefix(pdipc_handle* h, double* dst, const double* aos, int npts) {
  h->stage.assign(3 * (size_t)h->m.ld, 0.0);
  HIPCHK(hipMemcpyAsync(h->stage.data(), dst, 3 * (size_t)h->m.ld * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int i = 0; i < npts; i++)
    for (int k = 0; k < 3; k++) h->stage[(size_t)k * h->m.ld + i] = aos[3 * (size_t)i + k];
  HIPCHK(hipMemcpyAsync(dst, h->stage.data(), 3 * (size_t)h->m.ld * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return PDIPC_OK;
}
int get_soa(pdipc_handle* h, const double* src, double* aos, int npts) {
  h->stage.assign(3 * (size_t)h->m.ld, 0.0);
  HIPCHK(hipMemcpyAsync(h->stage.data(), src, 3 * (size_t)h->m.ld * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int i = 0; i < npts; i++)
    for (int k = 0; k < 3; k++) aos[3 * (size_t)i + k] = h->stage[(size_t)k * h->m.ld + i];
  return PDIPC_OK;
}

int check_launch(pdipc_handle* h) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(PDIPC_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  (void)h;
  return PDIPC_OK;
}

int read_scalars(pdipc_handle* h) {
  HIPCHK(hipMemcpyAsync(h->scalHost, h->scalDev, 64, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return PDIPC_OK;
}

int upload_all(pdipc_handle* h) {
  HostSetup& hs = h->hs;
  DeviceMesh& m = h->m;
  int lanes = h->prm.sell_lanes;
  if (lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8) lanes = 8;
  build_sell(hs, lanes, h->sell);
  m.nTotal = hs.nTotal;
  m.nSim = hs.nSim;
  m.nCloth = hs.nCloth;
  m.nTets = hs.nTets;
  m.ld = ((hs.nTotal + 63) / 64) * 64;
  size_t V3 = 3 * (size_t)m.ld;
  int rc;
#define UP(field, vec) if ((rc = h->upload(&m.field, vec))) return rc
#define AL(field, n) if ((rc = h->alloc(&m.field, n))) return rc
  AL(X, V3); AL(Xold, V3); AL(X0, V3); AL(Dir, V3); AL(Pre, V3); AL(Next, V3); AL(V, V3);
  AL(initRhs, V3); AL(sysRhs, V3); AL(firstTerm, V3); AL(colRhs, V3); AL(clothA, V3); AL(Xrest, V3);
  AL(dPre, m.ld); AL(dCur, m.ld); AL(dNext, m.ld); AL(r0, m.ld);
  AL(xa0, m.ld); AL(xa1, m.ld); AL(xa2, m.ld); AL(cA, m.ld);
  for (Row64* p : {m.xa0, m.xa1, m.xa2, m.cA}) HIPCHK(hipMemset(p, 0, m.ld * sizeof(Row64)));
  for (double* p : {m.X, m.Xold, m.X0, m.Dir, m.Pre, m.Next, m.V, m.initRhs, m.sysRhs, m.firstTerm, m.colRhs, m.clothA,
                    m.Xrest})
    HIPCHK(hipMemset(p, 0, V3 * sizeof(double)));
  for (float4* p : {m.dPre, m.dCur, m.dNext, m.r0}) HIPCHK(hipMemset(p, 0, m.ld * sizeof(float4)));
  UP(mass, hs.mass); UP(invMass, hs.invMass); UP(gravY, hs.gravY); UP(fixK, hs.fixK); UP(fixFlag, hs.fixFlag);
  UP(colMass, hs.colMass); UP(elasticDiag, hs.elasticDiag);
  AL(inertialDiag, hs.nSim); AL(sysDiag, hs.nSim); AL(colDiag, hs.nTotal);
  AL(dStale, hs.nSim); AL(dChanged, hs.nSim); AL(dDirty, hs.nSim);
  HIPCHK(hipMemset(m.dStale, 1, std::max(1, hs.nSim)));  // first update recomputes every row
  HIPCHK(hipMemset(m.sysDiag, 0, std::max(1, hs.nSim) * sizeof(double)));
  HIPCHK(hipMemset(m.inertialDiag, 0, std::max(1, hs.nSim) * sizeof(double)));
  HIPCHK(hipMemset(m.colDiag, 0, std::max(1, hs.nTotal) * sizeof(double)));
  UP(cCenter, hs.cCenter); UP(cOff, hs.cOff.empty() ? std::vector<int>{0} : hs.cOff); UP(cNbr, hs.cNbr);
  UP(clothOfPoint, hs.clothOfPoint); UP(cRest, hs.cRest); UP(cLbo, hs.cLbo); UP(cBendIn, hs.cBendIn);
  {
    std::vector<int2> rr(std::max(1, hs.nSim), make_int2(0, 0));
    for (int i = 0; i < hs.nSim; i++) {
      const int c = hs.clothOfPoint[i];
      if (c >= 0) rr[i] = make_int2(hs.cOff[c], hs.cOff[c + 1]);
    }
    UP(ringRange, rr);
  }
  UP(cStrain, hs.cStrain); UP(cH0, hs.cH0);
  {
    std::vector<int4> t(hs.nTets);
    for (int e = 0; e < hs.nTets; e++) t[e] = make_int4(hs.tetIdx[4 * e], hs.tetIdx[4 * e + 1], hs.tetIdx[4 * e + 2], hs.tetIdx[4 * e + 3]);
    UP(tetIdx, t);
    std::vector<double> dm(9 * (size_t)hs.nTets);
    for (int e = 0; e < hs.nTets; e++)
      for (int k = 0; k < 9; k++) dm[(size_t)k * hs.nTets + e] = hs.tetDminv[9 * (size_t)e + k];
    UP(tetDminv, dm);
  }
  UP(tetW, hs.tetW); AL(tetOut, 12 * (size_t)hs.nTets);
  UP(vtOff, hs.vtOff); UP(vtList, hs.vtList);
  UP(r1Off, hs.r1Off); UP(r1Col, hs.r1Col); UP(r1Val, hs.r1Val);
  const SellLayout& L = h->sell;
  m.sellLanes = L.lanes;
  m.sellUniformW = L.uniformW;
  m.rowsPerSlice = L.rowsPerSlice;
  m.nSlices = L.nSlices;
  m.sellEntries = L.sliceOff.back();
  UP(sliceOff, L.sliceOff); UP(sliceW, L.sliceW); UP(m2Col, L.col); UP(m2RowFlag, L.rowFlag); UP(m2R1, L.r1);
  if (L.col16ok) UP(m2Col16, L.col16);
  if (!L.perm.empty()) {
    UP(sellPerm, L.perm);
    UP(sellSlotOf, L.slotOf);
  }
  UP(p2Off, L.pOff); UP(p2K, L.pK); UP(p2A, L.pA);
  AL(m2Val, m.sellEntries); AL(m2ValF, m.sellEntries);
  if (L.stageOk && h->prm.sell_stage) {
    // the tiles' column lists at one stride (the largest union rounded up to
    // 8 entries), each padded with its last column
    const int nTiles = (int)L.tileBase.size();
    const int stride = (L.stageMax + 7) / 8 * 8;
    std::vector<uint16_t> padded((size_t)nTiles * stride);
    for (int t = 0; t < nTiles; t++) {
      const int u0 = L.tileOff[t], nU = L.tileOff[t + 1] - u0;
      for (int u = 0; u < stride; u++) padded[(size_t)t * stride + u] = L.tileCol[u0 + std::min(u, nU - 1)];
    }
    UP(tileBase, L.tileBase); UP(tileCol, padded); UP(m2Loc, L.loc);
    m.stageStride = stride;
    // narrow staging for the FP32 and FP64 sweeps wherever its tiles fit (the
    // one-byte indices save a sixth of an FP32 sweep's bytes, a tenth of an
    // FP64 one's; FP32 with the matrix in HBM 260 -> 233 us per sweep at
    // 3 x 1024^2 cloths; with it Infinity-Cache resident for a 256^2 cloth
    // 1.7% more FP32-split and 4.8% more FP64 steps/s); sell_stage = 2 keeps
    // the 16-bit tiles
    if (L.narrowOk && h->prm.sell_stage != 2) {
      const int nN = (int)L.ntileBase.size();
      const int nStride = (L.narrowMax + 7) / 8 * 8;
      std::vector<uint16_t> npad((size_t)nN * nStride);
      for (int t = 0; t < nN; t++) {
        const int u0 = L.ntileOff[t], nU = L.ntileOff[t + 1] - u0;
        for (int u = 0; u < nStride; u++) npad[(size_t)t * nStride + u] = L.ntileCol[u0 + std::min(u, nU - 1)];
      }
      UP(ntileBase, L.ntileBase); UP(ntileCol, npad); UP(m2Loc8, L.loc8);
      m.narrowStride = nStride;
    }
  }
  // collision
  m.nFaces = (int)hs.bFaces.size() / 3;
  m.nEdges = (int)hs.bEdges.size() / 2;
  m.nPatches = (int)hs.patchStatic.size();
  {
    std::vector<int3> f(m.nFaces);
    for (int i = 0; i < m.nFaces; i++) f[i] = make_int3(hs.bFaces[3 * i], hs.bFaces[3 * i + 1], hs.bFaces[3 * i + 2]);
    UP(bFaces, f);
    std::vector<int2> e(m.nEdges);
    for (int i = 0; i < m.nEdges; i++) e[i] = make_int2(hs.bEdges[2 * i], hs.bEdges[2 * i + 1]);
    UP(bEdges, e);
  }
  UP(pfOff, hs.pfOff); UP(pfList, hs.pfList); UP(ppOff, hs.ppOff); UP(ppList, hs.ppList); UP(peOff, hs.peOff);
  UP(peList, hs.peList); UP(patchStatic, hs.patchStatic);
  int P = std::max(1, m.nPatches);
  AL(patchBox, 6 * (size_t)P); AL(sceneBox, 6); AL(sceneBoxPart, 6 * 256);
  AL(mortonA, P); AL(mortonB, P); AL(idxA, P); AL(idxB, P); AL(keys, P);
  int nbRadix = (P + 2047) / 2048;
  AL(radixHist, 256 * (size_t)nbRadix); AL(radixBlockSums, (256 * (size_t)nbRadix + 1023) / 1024 + 2);
  AL(nodeLeft, P); AL(nodeRight, P); AL(nodeParent, 2 * (size_t)P); AL(nodeRange, P);
  AL(nodes, 2 * (size_t)P);
  {
    int levels = 1;
    while ((1 << levels) <= P) levels++;
    AL(nodeTable, (size_t)levels * 6 * P);
  }
  const int hint = h->prm.capacity_hint;  // > 0: every growable buffer starts at this size
  {
    int nw = traverse_top_waves(P);
    m.segCapI = hint > 0 ? hint : 256;
    m.segCapP = m.segCapI;
    AL(segI, (size_t)nw * m.segCapI); AL(segP, (size_t)nw * m.segCapP);
    AL(cntI, nw); AL(offI, nw); AL(cntP, nw); AL(offP, nw);
    m.itemCap = hint > 0 ? hint : std::max(1024, 2 * P);
    AL(items, m.itemCap);
    AL(segS, (size_t)m.itemCap * traverse_sub_cap()); AL(cntS, m.itemCap); AL(offS, m.itemCap);
  }
  m.pairCap = hint > 0 ? hint : std::max(1024, 8 * P);
  AL(pairs, m.pairCap); AL(recCount, 2 * (size_t)m.pairCap); AL(recOff, 2 * (size_t)m.pairCap);
  m.recCap = hint > 0 ? hint : std::max(4096, 32 * P);
  AL(recs, m.recCap);
  m.scanBlockCap = (std::max(std::max(P, m.pairCap), m.itemCap) + 1023) / 1024 + 2;
  AL(scanBlock, m.scanBlockCap);
  m.actCap = hint > 0 ? hint : 4096;
  AL(actRec, m.actCap); AL(actW, m.actCap); AL(actT, 12 * (size_t)m.actCap);
  AL(entKey, 4 * (size_t)m.actCap); AL(entA, 4 * (size_t)m.actCap); AL(entPt, 4 * (size_t)m.actCap);
  AL(entSorted, 4 * (size_t)m.actCap);
  AL(ptCount, hs.nSim + 1); AL(ptOff, hs.nSim + 1); AL(ptCursor, hs.nSim + 1);
  AL(ptScan, (hs.nSim + 1 + 1023) / 1024 + 2);
  m.nPartials = (std::max(1, hs.nSim) + 255) / 256;
  AL(partials, m.nPartials);
  AL(checkCount, 1);
  // scalars: residual (8), toiBits (8), counters (10 ints)
  void* sd = nullptr;
  HIPCHK(hipMalloc(&sd, 64));
  h->allocs.push_back(sd);
  HIPCHK(hipMemset(sd, 0, 64));
  h->scalDev = sd;
  m.residual = (double*)sd;
  m.toiBits = (unsigned long long*)((char*)sd + 8);
  m.counters = (int*)((char*)sd + 16);
  HIPCHK(hipHostMalloc(&h->scalHost, 64, hipHostMallocDefault));
  memset(h->scalHost, 0, 64);
  // initial positions
  {
    std::vector<double> soa(V3, 0.0);
    for (int i = 0; i < hs.nTotal; i++)
      for (int k = 0; k < 3; k++) soa[(size_t)k * m.ld + i] = hs.X[3 * (size_t)i + k];
    for (double* p : {m.X, m.Xold, m.X0, m.Pre, m.Next, m.Xrest})
      HIPCHK(hipMemcpy(p, soa.data(), V3 * sizeof(double), hipMemcpyHostToDevice));
  }
#undef UP
#undef AL
  return PDIPC_OK;
}


// gpuCcdCulling (PdIpcSimulator.cpp:1799-1869) on (x, dx)
int broad_phase(pdipc_handle* h, const double* x, const double* dx) {
  DeviceMesh& m = h->m;
  hipStream_t s = h->stream;
  h->lastPairs = 0;
  h->lastRecs = 0;
  if (m.nPatches <= 0) return PDIPC_OK;
  double infl = h->prm.inflation >= 0 ? h->prm.inflation : 0.5 * h->prm.dhat;
  HIPCHK(hipMemsetAsync(m.counters, 0, 10 * sizeof(int), s));
  launch_patch_boxes(m, x, dx, infl, s);
  launch_lbvh(m, s);
  int rc;
  int nItems = 0, nTop = 0;
  for (int attempt = 0;; attempt++) {
    launch_traverse_top(m, h->prm.skip_static_pairs, s);
    if ((rc = check_launch(h))) return rc;
    if ((rc = read_scalars(h))) return rc;
    int need = h->hCounters()[3];
    nItems = h->hCounters()[6];
    nTop = h->hCounters()[7];
    if (need <= 0) break;
    if (attempt >= 2) return fail(PDIPC_E_CAPACITY, "traversal segments did not fit after regrowth");
    int cap = need + need / 2 + 64;  // grow the pass-1 segments and search again
    size_t nw = (size_t)traverse_top_waves(m.nPatches);
    h->release(m.segI);
    h->release(m.segP);
    m.segI = m.segP = nullptr;
    if ((rc = h->alloc(&m.segI, nw * cap)) || (rc = h->alloc(&m.segP, nw * cap))) return rc;
    m.segCapI = m.segCapP = cap;
    HIPCHK(hipMemsetAsync(m.counters + 3, 0, sizeof(int), s));
  }
  if (nItems > m.itemCap) {
    int cap = nItems + nItems / 2 + 1024;
    for (void* p : {(void*)m.items, (void*)m.segS, (void*)m.cntS, (void*)m.offS}) h->release(p);
    if ((rc = h->alloc(&m.items, cap)) || (rc = h->alloc(&m.segS, (size_t)cap * traverse_sub_cap())) ||
        (rc = h->alloc(&m.cntS, cap)) || (rc = h->alloc(&m.offS, cap)))
      return rc;
    m.itemCap = cap;
    int sb = (cap + 1023) / 1024 + 2;
    if (sb > m.scanBlockCap) {
      h->release(m.scanBlock);
      if ((rc = h->alloc(&m.scanBlock, sb))) return rc;
      m.scanBlockCap = sb;
    }
  }
  launch_traverse_sub(m, h->prm.skip_static_pairs, nItems, s);
  if ((rc = check_launch(h))) return rc;
  if ((rc = read_scalars(h))) return rc;
  int64_t Q64 = (int64_t)nTop + (nItems > 0 ? h->hCounters()[0] : 0);
  if (Q64 > INT32_MAX / 2) return fail(PDIPC_E_CAPACITY, "patch pair count overflows int32");
  int Q = (int)Q64;
  if (Q > m.pairCap) {
    int cap = Q + Q / 2 + 1024;
    h->release(m.pairs);
    if ((rc = h->alloc(&m.pairs, cap))) return rc;
    h->release(m.recCount);
    if ((rc = h->alloc(&m.recCount, 2 * (size_t)cap))) return rc;
    h->release(m.recOff);
    if ((rc = h->alloc(&m.recOff, 2 * (size_t)cap))) return rc;
    int sb = (cap + 1023) / 1024 + 2;
    if (sb > m.scanBlockCap) {
      h->release(m.scanBlock);
      if ((rc = h->alloc(&m.scanBlock, sb))) return rc;
      m.scanBlockCap = sb;
    }
    m.pairCap = cap;
  }
  launch_pairs_compact(m, nItems, nTop, s);
  launch_rec_count(m, Q, s);
  if ((rc = check_launch(h))) return rc;
  if ((rc = read_scalars(h))) return rc;
  int64_t K64 = Q > 0 ? (int64_t)h->hCounters()[1] + h->hCounters()[8] : 0;
  int nPT = Q > 0 ? h->hCounters()[1] : 0;
  if (nPT < 0 || h->hCounters()[8] < 0 || K64 > INT32_MAX / 2)
    return fail(PDIPC_E_CAPACITY, "candidate count overflows int32");
  int K = (int)K64;
  if (K > m.recCap) {
    int cap = K + K / 2 + 4096;
    h->release(m.recs);
    if ((rc = h->alloc(&m.recs, cap))) return rc;
    m.recCap = cap;
  }
  launch_rec_fill(m, Q, nPT, K, s);
  if ((rc = check_launch(h))) return rc;
  h->lastPairs = Q;
  h->lastRecs = K;
  return PDIPC_OK;
}

// collision terms at m.X into colDiag / colRhs (simulated rows), growing the
// active-record list when a pass overflows it
int collision(pdipc_handle* h) {
  DeviceMesh& m = h->m;
  const pdipc_params& p = h->prm;
  hipStream_t s = h->stream;
  int rc;
  for (int attempt = 0; attempt < 2; attempt++) {
    launch_collision(m, h->lastRecs, p.dhat, p.kappa, p.dt, s);  // clears its counters first
    if ((rc = check_launch(h))) return rc;
    if ((rc = read_scalars(h))) return rc;
    if (!h->hCounters()[5]) return PDIPC_OK;
    int A = h->hCounters()[2];
    int cap = A + A / 2 + 4096;
    for (void* q : {(void*)m.actRec, (void*)m.actW, (void*)m.actT, (void*)m.entKey, (void*)m.entA, (void*)m.entPt,
                    (void*)m.entSorted})
      h->release(q);
    m.actCap = cap;
    if ((rc = h->alloc(&m.actRec, cap)) || (rc = h->alloc(&m.actW, cap)) || (rc = h->alloc(&m.actT, 12 * (size_t)cap)) ||
        (rc = h->alloc(&m.entKey, 4 * (size_t)cap)) || (rc = h->alloc(&m.entA, 4 * (size_t)cap)) ||
        (rc = h->alloc(&m.entPt, 4 * (size_t)cap)) || (rc = h->alloc(&m.entSorted, 4 * (size_t)cap)))
      return rc;
  }
  return fail(PDIPC_E_CAPACITY, "active collision list did not fit after growing");
}

void enqueue_local_global(pdipc_handle* h, int withCol) {
  DeviceMesh& m = h->m;
  const pdipc_params& p = h->prm;
  hipStream_t s = h->stream;
  launch_diag(m, withCol, s);
  launch_operators(m, p.omega, p.precision, s);
  for (int it = 0; it < p.lg_iters; it++) {
    bool fused = p.precision == PDIPC_FP32_SPLIT;  // k_rhs also runs the residual's first pass
    launch_local(m, p.precision, p.sigma_lo, p.sigma_hi, withCol, s, fused ? p.omega : -1.0);
    launch_solve(m, p.omega, p.rho, p.cheb_iters, p.precision, s, fused);
  }
}

// LocalGlobalIteration (PdIpcSimulator.cpp:1554-1622).  Its kernel sequence and
// arguments are identical every outer iteration (the solver's buffer rotation
// is deterministic), so with use_graphs it is captured once and replayed.
int local_global(pdipc_handle* h, int withCol) {
  if (!h->prm.use_graphs || h->m.timer) {  // kernel timing times each launch: no graph replay
    enqueue_local_global(h, withCol);
    return check_launch(h);
  }
  hipGraphExec_t& g = h->lgGraph[withCol ? 1 : 0][h->prm.precision == 1 ? 1 : 0];
  if (!g) {
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    enqueue_local_global(h, withCol);
    HIPCHK(hipStreamEndCapture(h->stream, &graph));
    hipError_t e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      g = nullptr;
      return fail(PDIPC_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
  }
  HIPCHK(hipGraphLaunch(g, h->stream));
  return check_launch(h);
}

int advance(pdipc_handle* h) {
  DeviceMesh& m = h->m;
  const pdipc_params& p = h->prm;
  hipStream_t s = h->stream;
  size_t V3 = 3 * (size_t)m.ld;
  int rc;
  launch_copy(m.X0, m.X, (int64_t)V3, s);
  launch_copy(m.Xold, m.X, (int64_t)V3, s);
  launch_predict(m, p.dt, s);
  bool activeCol = false;
  int outer = 0;
  double toiMin = 1.0;
  uint64_t argminHash = 0, activeHash = 0;  // diagnostics (track_toi_argmin), wrapping arithmetic
  h->cnt.lg_iters = 0;
  for (; outer < p.max_outer; outer++) {
    if ((rc = local_global(h, activeCol ? 1 : 0))) return rc;
    h->cnt.lg_iters += p.lg_iters;
    launch_direction(m, s);
    if (outer == 0) {
      hipEvent_t b0 = nullptr, b1 = nullptr;
      if (m.timer && m.timer->take(TC_BROAD, &b0, &b1)) HIPCHK(hipEventRecord(b0, s));
      if ((rc = broad_phase(h, m.Xold, m.Dir))) return rc;
      if (b1) HIPCHK(hipEventRecord(b1, s));
    }
    const bool track = p.track_toi_argmin && h->lastRecs > 0;  // (launch_direction reset the TOI minimum)
    if (track) {  // per-record TOIs, then the first record attaining the minimum
      if (h->trackCap < h->lastRecs) {
        h->release(h->trackToi);
        h->trackToi = nullptr;
        h->trackCap = 0;
        if ((rc = h->alloc(&h->trackToi, (size_t)h->lastRecs))) return rc;
        h->trackCap = h->lastRecs;
      }
      HIPCHK(hipMemsetAsync(m.counters + 4, 0x7f, sizeof(int), s));
      launch_ccd_records(m, m.recs, h->lastRecs, m.Xold, m.X, p.ccd_filter, p.ccd_iters, p.ccd_step, p.ccd_gap,
                         p.touch_policy, h->trackToi, m.counters + 4, s);
      launch_argmin_path(m, h->lastRecs, m.Xold, m.X, p.ccd_filter, p.ccd_iters, p.ccd_step, p.ccd_gap,
                         p.touch_policy, m.counters + 4, s);
    } else {
      launch_ccd(m, h->lastRecs, p.ccd_filter, p.ccd_iters, p.ccd_step, p.ccd_gap, p.touch_policy, s);
    }
    launch_clamp_residual(m, p.toi_scale, s);
    if ((rc = collision(h))) return rc;  // also reads the scalars back
    if (p.track_toi_argmin) {  // the record attaining the minimum TOI, its ACCD exit path, the active set
      const int am = track ? h->hCounters()[4] : 0x7f7f7f7f;
      const bool hit = am >= 0 && am < h->lastRecs;
      argminHash = argminHash * 1000003ull + (hit ? (uint64_t)am + 1 : 0);
      argminHash = argminHash * 1000003ull + (hit ? (uint64_t)(h->hCounters()[9] + 2) : 0);
      unsigned long long* hashDev = (unsigned long long*)(m.counters + 10);
      HIPCHK(hipMemsetAsync(hashDev, 0, sizeof(unsigned long long), s));
      launch_active_hash(m, hashDev, s);
      if ((rc = read_scalars(h))) return rc;
      uint64_t act;
      memcpy(&act, h->hCounters() + 10, sizeof act);
      activeHash = activeHash * 1000003ull + act;
    }
    activeCol = true;
    double res = *h->hResidual();
    double tmin;
    unsigned long long b = *h->hToi();
    memcpy(&tmin, &b, 8);
    double toi = tmin < 1.0 ? tmin * p.toi_scale : 1.0;
    if (toi < toiMin) toiMin = toi;
    h->cnt.residual = res;
    h->cnt.active_pairs = h->hCounters()[2];
    if (!std::isfinite(res)) return fail(PDIPC_E_NAN, "non-finite search direction");
    if (res < p.residual_tol && outer > 0) break;
    launch_copy(m.Xold, m.X, (int64_t)V3, s);
  }
  h->cnt.outer_iters = outer < p.max_outer ? outer + 1 : p.max_outer;
  h->cnt.toi_min = toiMin;
  h->cnt.toi_argmins = (long long)argminHash;
  h->cnt.active_sets = (long long)activeHash;
  h->cnt.patch_pairs = h->lastPairs;
  h->cnt.candidates = h->lastRecs;
  launch_velocity(m, p.dt, s);
  return check_launch(h);
}

int require_device(pdipc_handle* h) {
  if (!h) return fail(PDIPC_E_ARG, "null handle");
  if (!h->deviceReady) return fail(PDIPC_E_ARG, "pdipc_init has not run");
  return PDIPC_OK;
}

ModelInput make_input(int kind, const double* v, int nv, const int* el, int nel, int per, double s0, double s1,
                      double s2, double density, int patchDim, const int* fixed, int nfixed) {
  ModelInput m;
  m.kind = kind;
  m.verts.assign(v, v + 3 * (size_t)nv);
  m.elems.assign(el, el + per * (size_t)nel);
  m.s0 = s0; m.s1 = s1; m.s2 = s2; m.density = density; m.patchDim = patchDim;
  if (fixed && nfixed > 0) m.fixed.assign(fixed, fixed + nfixed);
  return m;
}

}  // namespace pdipc

// Full-state checkpoint: the step reads nothing from earlier steps but the
// positions and velocities, so (x, v) restart a run bit for bit.  File:
// "PDIPCST1", int32 version 1, int32 nSim, int32 nTotal, int32 0, uint64
// topology fingerprint (FNV-1a over the point counts and the ring, tet and R1
// patterns), then x (3 nTotal doubles, AoS) and v (3 nSim doubles, AoS).
namespace {
constexpr char kStateMagic[8] = {'P', 'D', 'I', 'P', 'C', 'S', 'T', '1'};
struct StateHeader {
  char magic[8];
  int32_t version, nSim, nTotal, pad;
  uint64_t topo;
};
template <class T>
void fnv(uint64_t& hsh, const T* p, size_t n) {
  const unsigned char* b = reinterpret_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n * sizeof(T); i++) {
    hsh ^= b[i];
    hsh *= 1099511628211ull;
  }
}
uint64_t topology_fingerprint(const HostSetup& hs) {
  uint64_t hsh = 1469598103934665603ull;
  const int counts[4] = {hs.nTotal, hs.nSim, hs.nCloth, hs.nTets};
  fnv(hsh, counts, 4);
  fnv(hsh, hs.cOff.data(), hs.cOff.size());
  fnv(hsh, hs.cNbr.data(), hs.cNbr.size());
  fnv(hsh, hs.tetIdx.data(), hs.tetIdx.size());
  fnv(hsh, hs.r1Off.data(), hs.r1Off.size());
  fnv(hsh, hs.r1Col.data(), hs.r1Col.size());
  return hsh;
}
}  // namespace

extern "C" {

void pdipc_default_params(pdipc_params* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->dt = 0.01; o->gravity = -9.8; o->dhat = 1e-3; o->kappa = 1e4;
  o->lg_iters = 10; o->cheb_iters = 15; o->max_outer = 30; o->residual_tol = 1e-5;
  o->omega = 2.0 / 3.0; o->rho = 0.9992; o->sigma_lo = 0.999; o->sigma_hi = 1.001;
  o->ccd_filter = 5e-3; o->toi_scale = 0.98; o->ccd_iters = 10; o->ccd_step = 0.9; o->ccd_gap = 0.1;
  o->static_mass = 1e4; o->inflation = -1.0; o->skip_static_pairs = 0; o->touch_policy = 1;
  o->precision = PDIPC_FP64; o->sell_lanes = 8; o->use_graphs = 0; o->sell_stage = 1;
}

const char* pdipc_last_error(void) { return pdipc::g_err.c_str(); }

int pdipc_create(const pdipc_params* params, pdipc_handle** out) {
  if (!out) return fail(PDIPC_E_ARG, "null output pointer");
  auto* h = new pdipc_handle();
  if (params) h->prm = *params;
  else pdipc_default_params(&h->prm);
  if (h->prm.dt <= 0 || h->prm.lg_iters < 1 || h->prm.cheb_iters < 1 || h->prm.max_outer < 1) {
    delete h;
    return fail(PDIPC_E_ARG, "invalid step parameters");
  }
  memset(&h->cnt, 0, sizeof(h->cnt));
  *out = h;
  return PDIPC_OK;
}

int pdipc_destroy(pdipc_handle* h) {
  delete h;
  return PDIPC_OK;
}

int pdipc_add_cloth(pdipc_handle* h, const double* verts, int nverts, const int* tris, int ntris, double s_strain,
                    double s_bend, double s_pos, double density, int patch_dim, const int* fixed, int nfixed) {
  if (!h || !verts || !tris || nverts <= 0 || ntris <= 0) return fail(PDIPC_E_ARG, "bad cloth arguments");
  if (h->hostReady) return fail(PDIPC_E_ARG, "models must be added before setup");
  h->sims.push_back(make_input(MODEL_CLOTH, verts, nverts, tris, ntris, 3, s_strain, s_bend, s_pos, density, patch_dim,
                               fixed, nfixed));
  return PDIPC_OK;
}

int pdipc_add_tet(pdipc_handle* h, const double* verts, int nverts, const int* tets, int ntets, double s_strain,
                  double s_volume, double s_pos, double density, int patch_dim, const int* fixed, int nfixed) {
  if (!h || !verts || !tets || nverts <= 0 || ntets <= 0) return fail(PDIPC_E_ARG, "bad tet arguments");
  if (h->hostReady) return fail(PDIPC_E_ARG, "models must be added before setup");
  h->sims.push_back(make_input(MODEL_TET, verts, nverts, tets, ntets, 4, s_strain, s_volume, s_pos, density, patch_dim,
                               fixed, nfixed));
  return PDIPC_OK;
}

int pdipc_add_static(pdipc_handle* h, const double* verts, int nverts, const int* tris, int ntris, int patch_dim) {
  if (!h || !verts || !tris || nverts <= 0 || ntris <= 0) return fail(PDIPC_E_ARG, "bad static arguments");
  if (h->hostReady) return fail(PDIPC_E_ARG, "models must be added before setup");
  h->statics.push_back(make_input(MODEL_STATIC, verts, nverts, tris, ntris, 3, 0, 0, 0, 1e5, patch_dim, nullptr, 0));
  return PDIPC_OK;
}

int pdipc_set_cloth_stiffness(pdipc_handle* h, int model, const double* k, int ntris) {
  if (!h || !k) return fail(PDIPC_E_ARG, "null argument");
  if (model < 0 || model >= (int)h->sims.size() || h->sims[model].kind != MODEL_CLOTH)
    return fail(PDIPC_E_ARG, "model is not a simulated cloth");
  if (ntris * 3 != (int)h->sims[model].elems.size()) return fail(PDIPC_E_ARG, "one stiffness per cloth triangle");
  for (int t = 0; t < ntris; t++)
    if (!(k[t] > 0) || !std::isfinite(k[t])) return fail(PDIPC_E_ARG, "stiffness must be positive and finite");
  std::vector<double> old = h->sims[model].triK;
  h->sims[model].triK.assign(k, k + ntris);
  if (!h->hostReady) return PDIPC_OK;
  // recompute the values; the patterns must not change (they size the device arrays)
  StepConstants c;
  c.dt = h->prm.dt;
  c.gravity = h->prm.gravity;
  c.dhat = h->prm.dhat;
  HostSetup ns;
  std::string err = build_host_setup(h->sims, h->statics, c, ns);
  if (!err.empty() || ns.r1Col != h->hs.r1Col || ns.m2Col != h->hs.m2Col || ns.p2K != h->hs.p2K ||
      ns.cNbr != h->hs.cNbr) {
    h->sims[model].triK = old;
    return fail(PDIPC_E_ARG, err.empty() ? "stiffness change alters the operator pattern" : err);
  }
  for (int i = ns.nSim; i < ns.nTotal; i++) ns.colMass[i] = h->prm.static_mass;
  h->hs = std::move(ns);
  if (!h->deviceReady) return PDIPC_OK;
  build_sell(h->hs, h->sell.lanes, h->sell);
  DeviceMesh& m = h->m;
  const HostSetup& hs = h->hs;
  HIPCHK(hipStreamSynchronize(h->stream));  // the uploads below go through the null stream
  HIPCHK(hipMemcpy(m.elasticDiag, hs.elasticDiag.data(), hs.elasticDiag.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m.r1Val, hs.r1Val.data(), hs.r1Val.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m.cStrain, hs.cStrain.data(), hs.cStrain.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m.m2R1, h->sell.r1.data(), h->sell.r1.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m.p2A, h->sell.pA.data(), h->sell.pA.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(m.dStale, 1, std::max(1, hs.nSim)));  // every M^2 row is recomputed next
  return pdipc::adjoint_refresh_values(h);
}

int pdipc_host_setup(pdipc_handle* h) {
  if (!h) return fail(PDIPC_E_ARG, "null handle");
  if (h->hostReady) return PDIPC_OK;
  if (h->sims.empty()) return fail(PDIPC_E_ARG, "no simulated model");
  StepConstants k;
  k.dt = h->prm.dt;
  k.gravity = h->prm.gravity;
  k.dhat = h->prm.dhat;
  std::string err;
  try {
    err = build_host_setup(h->sims, h->statics, k, h->hs);
  } catch (const std::exception& e) {
    err = std::string("host setup: ") + e.what();
  }
  if (!err.empty()) return fail(PDIPC_E_BADMESH, err);
  for (int i = h->hs.nSim; i < h->hs.nTotal; i++) h->hs.colMass[i] = h->prm.static_mass;
  h->hostReady = true;
  HostSetup& hs = h->hs;
  h->cnt.sim_points = hs.nSim;
  h->cnt.total_points = hs.nTotal;
  h->cnt.cloth_points = hs.nCloth;
  h->cnt.tets = hs.nTets;
  h->cnt.patches = (int)hs.patchStatic.size();
  h->cnt.boundary_faces = (int)hs.bFaces.size() / 3;
  h->cnt.boundary_edges = (int)hs.bEdges.size() / 2;
  h->cnt.nnz_r1 = hs.nnzR1;
  h->cnt.nnz_r2 = hs.nnzR2;
  h->cnt.n_2paths = hs.n2paths;
  return PDIPC_OK;
}

int pdipc_init(pdipc_handle* h) {
  int rc = pdipc_host_setup(h);
  if (rc) return rc;
  if (h->deviceReady) return PDIPC_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(PDIPC_E_NODEVICE, "no HIP device");
  HIPCHK(hipGetDevice(&h->device));
  if ((rc = pdipc::noalloc_setup(h->device))) return rc;
  HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  if ((rc = upload_all(h))) return rc;
  h->cnt.sell_entries = h->m.sellEntries;
  h->deviceReady = true;
  return PDIPC_OK;
}

// adds the recorded launches' durations to the timer's sums
static int fold_timer(pdipc_handle* h) {
  pdipc::KernelTimer* t = h->m.timer;
  if (!t || !t->used) return PDIPC_OK;
  HIPCHK(hipStreamSynchronize(h->stream));
  for (size_t i = 0; i < t->used; i++) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, t->pool[2 * i], t->pool[2 * i + 1]));
    t->us[t->cls[i]] += 1000.0 * ms;
    t->count[t->cls[i]]++;
  }
  t->used = 0;
  return PDIPC_OK;
}

int pdipc_step(pdipc_handle* h, int nsteps) {
  int rc = require_device(h);
  if (rc) return rc;
  for (int i = 0; i < nsteps; i++) {
    if ((rc = advance(h))) return rc;
    if ((rc = fold_timer(h))) return rc;
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return PDIPC_OK;
}

int pdipc_kernel_timing(pdipc_handle* h, int enable) {
  int rc = require_device(h);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  delete h->m.timer;  // a new timer starts from zero
  h->m.timer = nullptr;
  if (enable) h->m.timer = new pdipc::KernelTimer();
  return PDIPC_OK;
}

int pdipc_get_kernel_times(pdipc_handle* h, double* us, long long* counts, int nclasses) {
  int rc = require_device(h);
  if (rc) return rc;
  if (!us || !counts || nclasses < 0) return fail(PDIPC_E_ARG, "null output");
  if ((rc = fold_timer(h))) return rc;
  for (int c = 0; c < nclasses; c++) {
    us[c] = (h->m.timer && c < pdipc::TC_COUNT) ? h->m.timer->us[c] : 0.0;
    counts[c] = (h->m.timer && c < pdipc::TC_COUNT) ? h->m.timer->count[c] : 0;
  }
  return PDIPC_OK;
}

int pdipc_set_precision(pdipc_handle* h, int precision) {
  if (!h) return fail(PDIPC_E_ARG, "null handle");
  if (precision != PDIPC_FP64 && precision != PDIPC_FP32_SPLIT) return fail(PDIPC_E_ARG, "unknown precision");
  if (h->prm.precision != precision) h->drop_graphs();
  h->prm.precision = precision;
  return PDIPC_OK;
}

int pdipc_get_x(pdipc_handle* h, double* out, int npoints) {
  int rc = require_device(h);
  if (rc) return rc;
  if (!out || npoints < h->m.nTotal) return fail(PDIPC_E_ARG, "output holds fewer points than the scene");
  return get_soa(h, h->m.X, out, h->m.nTotal);
}

int pdipc_set_x(pdipc_handle* h, const double* in, int npoints) {
  int rc = require_device(h);
  if (rc) return rc;
  if (!in || npoints < h->m.nSim || npoints > h->m.nTotal) return fail(PDIPC_E_ARG, "npoints must cover the simulated points");
  return put_soa_prefix(h, h->m.X, in, h->m.nSim);
}

int pdipc_get_v(pdipc_handle* h, double* out, int npoints) {
  int rc = require_device(h);
  if (rc) return rc;
  if (!out || npoints < h->m.nSim) return fail(PDIPC_E_ARG, "output holds fewer points than the simulated set");
  return get_soa(h, h->m.V, out, h->m.nSim);
}

int pdipc_set_v(pdipc_handle* h, const double* in, int npoints) {
  int rc = require_device(h);
  if (rc) return rc;
  if (!in || npoints < h->m.nSim) return fail(PDIPC_E_ARG, "npoints must cover the simulated points");
  return put_soa(h, h->m.V, in, h->m.nSim);
}

int pdipc_get_counters(pdipc_handle* h, pdipc_counters* out) {
  if (!h || !out) return fail(PDIPC_E_ARG, "null argument");
  *out = h->cnt;
  return PDIPC_OK;
}

int pdipc_save_frame(pdipc_handle* h, const char* path, int as_double) {
  int rc = require_device(h);
  if (rc) return rc;
  std::vector<double> x(3 * (size_t)h->m.nTotal);
  if ((rc = get_soa(h, h->m.X, x.data(), h->m.nTotal))) return rc;
  FILE* f = fopen(path, "wb");
  if (!f) return fail(PDIPC_E_ARG, std::string("cannot open ") + path);
  int64_t rows = 3, cols = h->m.nTotal;
  fwrite(&rows, 8, 1, f);
  fwrite(&cols, 8, 1, f);
  if (as_double) fwrite(x.data(), sizeof(double), x.size(), f);
  else {
    std::vector<float> xf(x.begin(), x.end());
    fwrite(xf.data(), sizeof(float), xf.size(), f);
  }
  fclose(f);
  return PDIPC_OK;
}

int pdipc_save_state(pdipc_handle* h, const char* path) {
  int rc = require_device(h);
  if (rc) return rc;
  if (!path) return fail(PDIPC_E_ARG, "null path");
  const DeviceMesh& m = h->m;
  std::vector<double> x(3 * (size_t)m.nTotal), v(3 * (size_t)m.nSim);
  if ((rc = get_soa(h, m.X, x.data(), m.nTotal)) || (rc = get_soa(h, m.V, v.data(), m.nSim))) return rc;
  StateHeader hd;
  memcpy(hd.magic, kStateMagic, 8);
  hd.version = 1;
  hd.nSim = m.nSim;
  hd.nTotal = m.nTotal;
  hd.pad = 0;
  hd.topo = topology_fingerprint(h->hs);
  FILE* f = fopen(path, "wb");
  if (!f) return fail(PDIPC_E_ARG, std::string("cannot open ") + path);
  bool ok = fwrite(&hd, sizeof(hd), 1, f) == 1 && fwrite(x.data(), sizeof(double), x.size(), f) == x.size() &&
            fwrite(v.data(), sizeof(double), v.size(), f) == v.size();
  ok = (fclose(f) == 0) && ok;
  return ok ? PDIPC_OK : fail(PDIPC_E_ARG, std::string("short write to ") + path);
}

int pdipc_load_state(pdipc_handle* h, const char* path) {
  int rc = require_device(h);
  if (rc) return rc;
  if (!path) return fail(PDIPC_E_ARG, "null path");
  const DeviceMesh& m = h->m;
  FILE* f = fopen(path, "rb");
  if (!f) return fail(PDIPC_E_ARG, std::string("cannot open ") + path);
  StateHeader hd;
  std::vector<double> x(3 * (size_t)m.nTotal), v(3 * (size_t)m.nSim);
  bool ok = fread(&hd, sizeof(hd), 1, f) == 1;
  std::string why;
  if (!ok || memcmp(hd.magic, kStateMagic, 8) != 0 || hd.version != 1)
    why = "not a state file";
  else if (hd.nSim != m.nSim || hd.nTotal != m.nTotal || hd.topo != topology_fingerprint(h->hs))
    why = "saved from a different scene";
  else if (fread(x.data(), sizeof(double), x.size(), f) != x.size() ||
           fread(v.data(), sizeof(double), v.size(), f) != v.size())
    why = "truncated";
  else {
    char extra;
    if (fread(&extra, 1, 1, f) != 0) why = "trailing data";
  }
  fclose(f);
  if (!why.empty()) return fail(PDIPC_E_ARG, std::string(path) + ": " + why);
  if ((rc = put_soa(h, h->m.X, x.data(), m.nTotal)) || (rc = put_soa(h, h->m.V, v.data(), m.nSim))) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  return PDIPC_OK;
}

int pdipc_get_candidates(pdipc_handle* h, int* out, int cap, int* count) {
  int rc = require_device(h);
  if (rc) return rc;
  int K = h->lastRecs;
  if (count) *count = K;
  if (!out || cap <= 0 || K == 0) return PDIPC_OK;
  int n = std::min(K, cap);
  std::vector<int4> r(n);
  HIPCHK(hipMemcpyAsync(r.data(), h->m.recs, n * sizeof(int4), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int i = 0; i < n; i++) {
    bool ee = r[i].x < 0;
    out[5 * i] = ee ? 1 : 0;
    out[5 * i + 1] = ee ? ~r[i].x : r[i].x;
    out[5 * i + 2] = r[i].y;
    out[5 * i + 3] = r[i].z;
    out[5 * i + 4] = r[i].w;
  }
  return PDIPC_OK;
}

int pdipc_get_patch_pairs(pdipc_handle* h, int* out, int cap, int* count) {
  int rc = require_device(h);
  if (rc) return rc;
  int Q = h->lastPairs;
  if (count) *count = Q;
  if (!out || cap <= 0 || Q == 0) return PDIPC_OK;
  int n = std::min(Q, cap);
  HIPCHK(hipMemcpyAsync(out, h->m.pairs, n * sizeof(int2), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return PDIPC_OK;
}

int pdipc_get_model_dat(pdipc_handle* h, int model, int which, int* out, int cap, int* count) {
  if (!h) return fail(PDIPC_E_ARG, "null handle");
  if (!h->hostReady) return fail(PDIPC_E_ARG, "pdipc_host_setup has not run");
  if (model < 0 || model >= (int)h->hs.boundaryDat.size()) return fail(PDIPC_E_ARG, "model index out of range");
  const std::vector<int>& d = which == 0 ? h->hs.boundaryDat[model] : h->hs.patchDat[model];
  if (count) *count = (int)d.size();
  if (out && cap > 0) std::copy(d.begin(), d.begin() + std::min<size_t>(cap, d.size()), out);
  return PDIPC_OK;
}

int pdipc_debug_sell_layout(pdipc_handle* h, int lanes, long long* sizes, long long* slice_off, int* slice_w,
                            int* col, int* row_flag, short* col16, int* perm) {
  if (!h) return fail(PDIPC_E_ARG, "null handle");
  if (!h->hostReady) return fail(PDIPC_E_ARG, "pdipc_host_setup has not run");
  if (lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8) return fail(PDIPC_E_ARG, "lanes must be 1, 2, 4 or 8");
  if (!sizes) return fail(PDIPC_E_ARG, "null sizes");
  SellLayout L;
  build_sell(h->hs, lanes, L);
  const int64_t nE = L.sliceOff.back();
  sizes[0] = L.nSlices;
  sizes[1] = nE;
  sizes[2] = L.uniformW;
  sizes[3] = L.col16ok ? 1 : 0;
  sizes[4] = (long long)L.perm.size();
  sizes[5] = L.rowsPerSlice;
  if (slice_off) std::copy(L.sliceOff.begin(), L.sliceOff.end(), slice_off);
  if (slice_w) std::copy(L.sliceW.begin(), L.sliceW.end(), slice_w);
  if (col) std::copy(L.col.begin(), L.col.begin() + nE, col);
  if (row_flag) std::copy(L.rowFlag.begin(), L.rowFlag.begin() + nE, row_flag);
  if (col16 && L.col16ok) std::copy(L.col16.begin(), L.col16.begin() + nE, col16);
  if (perm) std::copy(L.perm.begin(), L.perm.end(), perm);
  return PDIPC_OK;
}

int pdipc_debug_stage_layout(pdipc_handle* h, int lanes, long long* sizes, int* tile_off, int* tile_base,
                             unsigned short* tile_col, unsigned short* loc) {
  if (!h) return fail(PDIPC_E_ARG, "null handle");
  if (!h->hostReady) return fail(PDIPC_E_ARG, "pdipc_host_setup has not run");
  if (lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8) return fail(PDIPC_E_ARG, "lanes must be 1, 2, 4 or 8");
  if (!sizes) return fail(PDIPC_E_ARG, "null sizes");
  SellLayout L;
  build_sell(h->hs, lanes, L);
  const int nTiles = L.stageOk ? (int)L.tileBase.size() : 0;
  sizes[0] = L.stageOk ? 1 : 0;
  sizes[1] = nTiles;
  sizes[2] = L.stageOk ? (long long)L.tileCol.size() : 0;
  sizes[3] = L.stageMax;
  sizes[4] = L.sliceOff.back();
  sizes[5] = kTileSlices;
  if (!L.stageOk) return PDIPC_OK;
  if (tile_off) std::copy(L.tileOff.begin(), L.tileOff.end(), tile_off);
  if (tile_base) std::copy(L.tileBase.begin(), L.tileBase.end(), tile_base);
  if (tile_col) std::copy(L.tileCol.begin(), L.tileCol.end(), tile_col);
  if (loc) std::copy(L.loc.begin(), L.loc.end(), loc);
  return PDIPC_OK;
}

int pdipc_debug_narrow_stage_layout(pdipc_handle* h, long long* sizes, int* tile_off, int* tile_base,
                                    unsigned short* tile_col, unsigned char* loc) {
  if (!h) return fail(PDIPC_E_ARG, "null handle");
  if (!h->hostReady) return fail(PDIPC_E_ARG, "pdipc_host_setup has not run");
  if (!sizes) return fail(PDIPC_E_ARG, "null sizes");
  SellLayout L;
  build_sell(h->hs, 8, L);
  const int nTiles = L.narrowOk ? (int)L.ntileBase.size() : 0;
  sizes[0] = L.narrowOk ? 1 : 0;
  sizes[1] = nTiles;
  sizes[2] = L.narrowOk ? (long long)L.ntileCol.size() : 0;
  sizes[3] = L.narrowMax;
  sizes[4] = L.sliceOff.back();
  sizes[5] = kNarrowTileSlices;
  if (!L.narrowOk) return PDIPC_OK;
  if (tile_off) std::copy(L.ntileOff.begin(), L.ntileOff.end(), tile_off);
  if (tile_base) std::copy(L.ntileBase.begin(), L.ntileBase.end(), tile_base);
  if (tile_col) std::copy(L.ntileCol.begin(), L.ntileCol.end(), tile_col);
  if (loc) std::copy(L.loc8.begin(), L.loc8.end(), loc);
  return PDIPC_OK;
}

int pdipc_sweep_staging(pdipc_handle* h, int* path) {
  int rc = require_device(h);
  if (rc) return rc;
  if (!path) return fail(PDIPC_E_ARG, "null path");
  *path = h->m.m2Loc8 ? 2 : (h->m.m2Loc ? 1 : 0);
  return PDIPC_OK;
}

int pdipc_debug_broad_phase(pdipc_handle* h, const double* xold, const double* dx, int* count) {
  int rc = require_device(h);
  if (rc) return rc;
  if ((rc = put_soa(h, h->m.Xold, xold, h->m.nTotal))) return rc;
  std::vector<double> d(3 * (size_t)h->m.nTotal, 0.0);
  std::copy(dx, dx + 3 * (size_t)h->m.nSim, d.begin());  // static points never move
  if ((rc = put_soa(h, h->m.Dir, d.data(), h->m.nTotal))) return rc;
  if ((rc = broad_phase(h, h->m.Xold, h->m.Dir))) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  if (count) *count = h->lastRecs;
  return PDIPC_OK;
}

int pdipc_debug_toi(pdipc_handle* h, const double* x0, const double* x1, double* toi) {
  int rc = require_device(h);
  if (rc) return rc;
  if ((rc = put_soa(h, h->m.Xold, x0, h->m.nTotal))) return rc;
  if ((rc = put_soa(h, h->m.X, x1, h->m.nTotal))) return rc;
  launch_toi_init(h->m, h->stream);
  launch_ccd(h->m, h->lastRecs, h->prm.ccd_filter, h->prm.ccd_iters, h->prm.ccd_step, h->prm.ccd_gap,
             h->prm.touch_policy, h->stream);
  if ((rc = check_launch(h))) return rc;
  if ((rc = read_scalars(h))) return rc;
  double tmin;
  unsigned long long b = *h->hToi();
  memcpy(&tmin, &b, 8);
  *toi = tmin < 1.0 ? tmin * h->prm.toi_scale : 1.0;
  return PDIPC_OK;
}

int pdipc_debug_collision(pdipc_handle* h, const double* x, double* diag, double* rhs, int* active) {
  int rc = require_device(h);
  if (rc) return rc;
  DeviceMesh& m = h->m;
  if ((rc = put_soa(h, m.X, x, m.nTotal))) return rc;
  if ((rc = collision(h))) return rc;
  if (active) *active = h->hCounters()[2];
  HIPCHK(hipMemcpyAsync(diag, m.colDiag, m.nSim * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  if ((rc = get_soa(h, m.colRhs, rhs, m.nSim))) return rc;
  return PDIPC_OK;
}

int pdipc_debug_local(pdipc_handle* h, const double* x, const double* init_rhs, double* rhs) {
  int rc = require_device(h);
  if (rc) return rc;
  DeviceMesh& m = h->m;
  if ((rc = put_soa(h, m.X, x, m.nTotal))) return rc;
  if ((rc = put_soa(h, m.initRhs, init_rhs, m.nSim))) return rc;
  launch_local(m, h->prm.precision, h->prm.sigma_lo, h->prm.sigma_hi, 0, h->stream);
  if ((rc = check_launch(h))) return rc;
  return get_soa(h, m.sysRhs, rhs, m.nSim);
}

int pdipc_debug_solve(pdipc_handle* h, const double* diag, const double* rhs, const double* x0, double* xout) {
  int rc = require_device(h);
  if (rc) return rc;
  DeviceMesh& m = h->m;
  HIPCHK(hipMemcpyAsync(m.sysDiag, diag, m.nSim * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemsetAsync(m.dChanged, 1, m.nSim, h->stream));  // the host wrote D: every row is stale
  HIPCHK(hipMemsetAsync(m.dStale, 1, m.nSim, h->stream));    // and stays stale for the next step
  if ((rc = put_soa(h, m.sysRhs, rhs, m.nSim))) return rc;
  if ((rc = put_soa(h, m.X, x0, m.nTotal))) return rc;
  launch_operators(m, h->prm.omega, h->prm.precision, h->stream);
  launch_solve(m, h->prm.omega, h->prm.rho, h->prm.cheb_iters, h->prm.precision, h->stream);
  if ((rc = check_launch(h))) return rc;
  return get_soa(h, m.X, xout, m.nSim);
}

int pdipc_count_intersections(pdipc_handle* h, long long* count) {
  int rc = require_device(h);
  if (rc) return rc;
  if (!count) return fail(PDIPC_E_ARG, "count is null");
  DeviceMesh& m = h->m;
  hipStream_t s = h->stream;
  launch_zero(m.Dir, 3 * (int64_t)m.ld, s);  // boxes of the current positions, no sweep
  if ((rc = broad_phase(h, m.X, m.Dir))) return rc;
  HIPCHK(hipMemsetAsync(m.checkCount, 0, sizeof(unsigned long long), s));
  launch_count_crossings(m, h->lastPairs, m.checkCount, s);
  if ((rc = check_launch(h))) return rc;
  unsigned long long c = 0;
  HIPCHK(hipMemcpyAsync(&c, m.checkCount, sizeof(c), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  *count = (long long)c;
  return PDIPC_OK;
}

int pdipc_debug_lbvh(pdipc_handle* h, unsigned long long* keys, int* left, int* right, int* parent, int* npatches) {
  int rc = require_device(h);
  if (rc) return rc;
  int P = h->m.nPatches;
  if (npatches) *npatches = P;
  if (P <= 0) return PDIPC_OK;
  if (keys) HIPCHK(hipMemcpyAsync(keys, h->m.keys, P * sizeof(uint64_t), hipMemcpyDeviceToHost, h->stream));
  if (P > 1) {
    if (left) HIPCHK(hipMemcpyAsync(left, h->m.nodeLeft, (P - 1) * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    if (right) HIPCHK(hipMemcpyAsync(right, h->m.nodeRight, (P - 1) * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    if (parent) HIPCHK(hipMemcpyAsync(parent, h->m.nodeParent, (2 * P - 1) * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return PDIPC_OK;
}

int pdipc_time_kernel(pdipc_handle* h, int which, int reps, double* avg_us, double* bytes) {
  int rc = require_device(h);
  if (rc) return rc;
  if (reps < 1 || !avg_us) return fail(PDIPC_E_ARG, "reps must be positive");
  DeviceMesh& m = h->m;
  const pdipc_params& p = h->prm;
  hipStream_t s = h->stream;
  auto launch = [&]() {
    switch (which) {
      case 0: launch_one_sweep(m, 0, s); break;
      case 1: launch_one_sweep(m, 1, s); break;
      case 2: launch_local(m, p.precision, p.sigma_lo, p.sigma_hi, 0, s); break;
      case 3:
        launch_toi_init(m, s);
        launch_ccd(m, h->lastRecs, p.ccd_filter, p.ccd_iters, p.ccd_step, p.ccd_gap, p.touch_policy, s);
        break;
      default: launch_operators(m, p.omega, p.precision, s); break;
    }
  };
  if (which < 0 || which > 4) return fail(PDIPC_E_ARG, "unknown kernel id");
  double N = m.nSim, nnz = (double)h->hs.nnzR2, D3 = 3.0 * N;
  double b = 0;
  switch (which) {
    case 0: b = nnz * (m.m2Col16 ? 10.0 : 12.0) + D3 * 8.0 * 4.0; break;  // col (16/32-bit)+val FP64; x_k, x_{k-1}, c, x_{k+1}
    case 1: b = nnz * (m.m2Col16 ? 6.0 : 8.0) + D3 * 4.0 * 4.0; break;  // col (16/32-bit)+val FP32; d_k, d_{k-1}, r0, d_{k+1}
    case 2: {
      double slots = (double)h->hs.cNbr.size();
      b = slots * (4 + 8 + 8 + 8) + h->hs.nCloth * 3 * 8.0 * 2 + D3 * 8.0 * 4;  // ring data; a; x, rhs0, rhs
      b += h->hs.nTets * (16 + 72 + 8 + 96.0);
      break;
    }
    case 3: b = h->lastRecs * (16.0 + 24 * 8.0); break;          // record + x0, x1 of 4 points
    default: b = (double)h->hs.n2paths * 12.0 + (double)m.sellEntries * (4 + 8 + 16 + 8 + 4); break;
  }
  if (which == 4) HIPCHK(hipMemsetAsync(m.dChanged, 1, m.nSim, s));  // time the full recompute
  launch();  // warm
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, s));
  for (int i = 0; i < reps; i++) launch();
  HIPCHK(hipEventRecord(e1, s));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if ((rc = check_launch(h))) return rc;
  *avg_us = 1000.0 * ms / reps;
  if (bytes) *bytes = b;
  return PDIPC_OK;
}

}  // extern "C"
