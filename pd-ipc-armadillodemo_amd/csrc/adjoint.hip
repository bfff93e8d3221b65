// Reverse mode of the PD-IPC step (an extension: the reference has no
// derivative).  The step is differentiated as computed, in FP64:
//
//   * the step count, the outer-iteration count of every step, the candidate
//     records, the set of active barrier pairs and the record attaining the
//     minimum TOI are frozen (piecewise-constant choices; the derivative is
//     exact wherever they do not switch);
//   * the Chebyshev solve is reversed sweep by sweep on the same SELL matrix:
//     with u = D^{-1} xbar, M^T xbar = D M u (R1 is symmetric), so the adjoint
//     sweep is the forward SpMV shape; the M^2 entries' adjoints are gathered
//     on the fly and pulled back through the operator update to D and R1;
//   * the local step is reversed per element (spring Jacobians, the bending
//     rescale, the SVD strain projection's differential), the barrier terms
//     and the ACCD time of impact by forward-mode dual numbers over the four
//     points of a record.
//
// Memory: pdipc_rollout keeps (x, v) every `checkpoint_every` steps;
// pdipc_backward re-runs one segment at a time recording every Chebyshev
// iterate (15 x 10 x 3N doubles per outer iteration), then walks it backwards.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "handle.hpp"

namespace pdipc {

namespace {

// bump allocator over large device chunks, reset per segment
struct Pool {
  struct Chunk {
    char* p;
    size_t size;
  };
  std::vector<Chunk> chunks;
  size_t ci = 0, off = 0;
  int alloc(size_t bytes, void** out) {
    bytes = (bytes + 255) & ~(size_t)255;
    while (ci < chunks.size() && off + bytes > chunks[ci].size) {
      ci++;
      off = 0;
    }
    if (ci == chunks.size()) {
      size_t sz = std::max(bytes, (size_t)1 << 29);
      char* p = nullptr;
      HIPCHK(hipMalloc((void**)&p, sz));
      chunks.push_back({p, sz});
      off = 0;
    }
    *out = chunks[ci].p + off;
    off += bytes;
    return PDIPC_OK;
  }
  void reset() { ci = 0; off = 0; }
  void release() {
    for (auto& c : chunks) (void)hipFree(c.p);
    chunks.clear();
    reset();
  }
};

// The vectors one outer iteration records, laid out in one contiguous stage:
//   [0] Xold, [1] X when the local/global loop starts, [2, 2+LG) the rhs of
//   each pass, then the Chebyshev iterates 1..K of each pass, then D (nsim).
// The recorded forward and the reverse pass both work on a fixed stage (so
// their loops can be captured once in HIP graphs); one device copy per outer
// iteration moves the stage to or from the tape.
struct StageLayout {
  size_t V3 = 0, nSim = 0;
  int LG = 0, K = 0;
  size_t slots() const { return 2 + (size_t)LG + (size_t)LG * K; }
  size_t bytes() const { return slots() * V3 * sizeof(double) + ((nSim * sizeof(double) + 255) & ~(size_t)255); }
  double* xold(double* b) const { return b; }
  double* xstart(double* b) const { return b + V3; }
  double* rhs(double* b, int l) const { return b + (2 + (size_t)l) * V3; }
  // k-th iterate of pass l; k = 0 is the pass's input.  Iterates 1 .. K of a
  // pass are consecutive vectors (the reverse pass packs them by that stride).
  double* it(double* b, int l, int k) const {
    if (k == 0) return l == 0 ? xstart(b) : it(b, l - 1, K);
    return b + (2 + (size_t)LG + (size_t)l * K + (k - 1)) * V3;
  }
  double* D(double* b) const { return b + slots() * V3; }
};

struct OuterTape {
  double* blob = nullptr;  // this outer iteration's stage, on the tape
  int* act = nullptr;      // records active in the collision pass that ends it
  int nAct = 0;
  double toi = 1.0;
  int clamped = 0;
  int argmin = -1;
  int withCol = 0;
};

struct StepTape {
  double *x0 = nullptr, *v0 = nullptr;
  int4* recs = nullptr;
  int nRecs = 0;
  std::vector<OuterTape> outers;
};

}  // namespace

struct AdjointState {
  AdjMesh am;
  std::vector<void*> allocs;
  // 2-path term behind each midA entry: midA holds the SELL pA values, which
  // scale with the spring stiffness, so a stiffness change re-gathers them
  std::vector<int64_t> midQ;
  int nSteps = 0, every = 1;
  bool recorded = false;  // the rollout kept the tape of every step (every == 0)
  bool lastKept = false;  // seg holds the last checkpoint segment, recorded by the rollout
  std::vector<double*> ckX, ckV;
  double *endX = nullptr, *endV = nullptr;
  Pool tape;
  std::vector<StepTape> seg;
  double *ax = nullptr, *axo = nullptr, *asol = nullptr, *uA = nullptr, *uB = nullptr, *uC = nullptr, *uc = nullptr;
  double *bbar = nullptr, *aA = nullptr, *tetAdj = nullptr, *rhs0bar = nullptr, *axn = nullptr;
  double* colR[2] = {nullptr, nullptr};
  double* colD[2] = {nullptr, nullptr};
  double *xbarA = nullptr, *vbarA = nullptr, *xbarB = nullptr, *vbarB = nullptr;
  double *Dbar = nullptr, *DbarEl = nullptr, *Gbar = nullptr, *R1bar = nullptr, *sbar = nullptr;
  double* recToi = nullptr;
  int recToiCap = 0;
  double *partials = nullptr, *gscale = nullptr;
  StageLayout lay;
  double* stage = nullptr;
  std::vector<double*> uS;  // reverse Chebyshev iterates u_0 .. u_{K+1} (u_{K+1} stays zero)
  double* colRacc = nullptr;
  // local/global loops of the recorded forward and of the reverse pass,
  // captured once per collision flag
  hipGraphExec_t fwdGraph[2] = {nullptr, nullptr};
  hipGraphExec_t revGraph[2] = {nullptr, nullptr};
  double stats[4] = {0, 0, 0, 0};  // recorded outer iterations, tape bytes, segments, steps
};

void adjoint_free(AdjointState* a) {
  if (!a) return;
  for (auto& g : a->fwdGraph)
    if (g) (void)hipGraphExecDestroy(g);
  for (auto& g : a->revGraph)
    if (g) (void)hipGraphExecDestroy(g);
  for (void* p : a->allocs) (void)hipFree(p);
  for (double* p : a->ckX) (void)hipFree(p);
  for (double* p : a->ckV) (void)hipFree(p);
  a->tape.release();
  delete a;
}

namespace {

template <class T>
int dalloc(AdjointState* A, T** p, size_t n) {
  HIPCHK(hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)));
  A->allocs.push_back(*p);
  HIPCHK(hipMemset(*p, 0, std::max<size_t>(n, 1) * sizeof(T)));
  return PDIPC_OK;
}
template <class T>
int dupload(AdjointState* A, T** p, const std::vector<T>& v) {
  int rc = dalloc(A, p, v.size());
  if (rc) return rc;
  if (!v.empty()) HIPCHK(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return PDIPC_OK;
}

int ensure_state(pdipc_handle* h) {
  if (h->adj) return PDIPC_OK;
  auto* A = new AdjointState();
  h->adj = A;
  const HostSetup& hs = h->hs;
  const DeviceMesh& m = h->m;
  size_t V3 = 3 * (size_t)m.ld;
  int rc;
  if ((rc = dupload(A, &A->am.sellR1Idx, h->sell.r1Idx)) || (rc = dupload(A, &A->am.pIa, h->sell.pIa)) ||
      (rc = dupload(A, &A->am.pIb, h->sell.pIb)) || (rc = dupload(A, &A->am.cSlotR1, hs.cSlotR1)) ||
      (rc = dupload(A, &A->am.cBendOut, hs.cBendOut)) || (rc = dupload(A, &A->am.cLboIn, hs.cLboIn)))
    return rc;
  // 2-paths grouped by middle point, and by the spring R1 entries they use
  {
    const SellLayout& L = h->sell;
    const int64_t nE = (int64_t)L.pOff.size() - 1;
    const int n = hs.nSim;
    std::vector<int> midOff(n + 1, 0);
    for (int64_t e = 0; e < nE; e++)
      for (int64_t q = L.pOff[e]; q < L.pOff[e + 1]; q++) midOff[L.pK[q] + 1]++;
    for (int k = 0; k < n; k++) midOff[k + 1] += midOff[k];
    std::vector<int64_t> midEnt(midOff[n]);
    std::vector<double> midA(midOff[n]);
    A->midQ.assign(midOff[n], 0);
    std::vector<int> fill(midOff.begin(), midOff.end() - 1);
    for (int64_t e = 0; e < nE; e++)
      for (int64_t q = L.pOff[e]; q < L.pOff[e + 1]; q++) {
        int at = fill[L.pK[q]]++;
        midEnt[at] = e;
        midA[at] = L.pA[q];
        A->midQ[at] = q;
      }
    std::vector<int> qt(hs.r1Col.size(), -1), springQ;
    for (int q : hs.cSlotR1)
      if (q >= 0 && qt[q] < 0) { qt[q] = 0; springQ.push_back(q); }
    std::sort(springQ.begin(), springQ.end());
    for (size_t t = 0; t < springQ.size(); t++) qt[springQ[t]] = (int)t;
    std::vector<int> sOff(springQ.size() + 1, 0);
    for (int64_t e = 0; e < nE; e++)
      for (int64_t q = L.pOff[e]; q < L.pOff[e + 1]; q++) {
        if (qt[L.pIa[q]] >= 0) sOff[qt[L.pIa[q]] + 1]++;
        if (qt[L.pIb[q]] >= 0) sOff[qt[L.pIb[q]] + 1]++;
      }
    for (size_t t = 0; t < springQ.size(); t++) sOff[t + 1] += sOff[t];
    std::vector<int64_t> sEnt(sOff.back());
    std::vector<int> sMid(sOff.back()), sOther(sOff.back());
    std::vector<int> sf(sOff.begin(), sOff.end() - 1);
    for (int64_t e = 0; e < nE; e++)
      for (int64_t q = L.pOff[e]; q < L.pOff[e + 1]; q++) {
        int ia = L.pIa[q], ib = L.pIb[q];
        if (qt[ia] >= 0) { int at = sf[qt[ia]]++; sEnt[at] = e; sMid[at] = L.pK[q]; sOther[at] = ib; }
        if (qt[ib] >= 0) { int at = sf[qt[ib]]++; sEnt[at] = e; sMid[at] = L.pK[q]; sOther[at] = ia; }
      }
    A->am.nSpringQ = (int)springQ.size();
    if ((rc = dupload(A, &A->am.midOff, midOff)) || (rc = dupload(A, &A->am.midEnt, midEnt)) ||
        (rc = dupload(A, &A->am.midA, midA)) || (rc = dupload(A, &A->am.springQ, springQ)) ||
        (rc = dupload(A, &A->am.springOff, sOff)) || (rc = dupload(A, &A->am.springEnt, sEnt)) ||
        (rc = dupload(A, &A->am.springMid, sMid)) || (rc = dupload(A, &A->am.springOther, sOther)))
      return rc;
  }
  // per slice: the sorted union of the columns of its entries, and per union
  // column and row the entry's offset in the slice's SELL block
  {
    const SellLayout& L = h->sell;
    const int R = 64 / L.lanes, n = m.nSim;
    std::vector<int> uOff(L.nSlices + 1, 0), uCol, cols;
    std::vector<uint16_t> uPos;
    auto real = [&](int64_t e, int r) {  // padding slots repeat the row's own index unflagged
      return r < n && (L.col[e] != r || (L.rowFlag[e] & (int)0x80000000) != 0);
    };
    for (int sl = 0; sl < L.nSlices; sl++) {
      const int64_t e0 = L.sliceOff[sl], e1 = L.sliceOff[sl + 1];
      if (e1 - e0 > 65535) return fail(PDIPC_E_ARG, "adjoint: a SELL slice holds more than 65535 entries");
      cols.clear();
      for (int64_t e = e0; e < e1; e++) {
        const int slot = sl * R + L.lane_of(sl, e) / L.lanes;
        const int r = L.perm.empty() ? slot : L.perm[slot];
        if (real(e, r)) cols.push_back(L.col[e]);
      }
      std::sort(cols.begin(), cols.end());
      cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
      const size_t p0 = uPos.size();
      uPos.resize(p0 + cols.size() * R, (uint16_t)0xFFFF);
      for (int64_t e = e0; e < e1; e++) {
        const int slot = sl * R + L.lane_of(sl, e) / L.lanes;
        const int r = L.perm.empty() ? slot : L.perm[slot];
        if (!real(e, r)) continue;
        const size_t t = std::lower_bound(cols.begin(), cols.end(), L.col[e]) - cols.begin();
        uPos[p0 + t * R + (slot - sl * R)] = (uint16_t)(e - e0);
      }
      uCol.insert(uCol.end(), cols.begin(), cols.end());
      uOff[sl + 1] = (int)uCol.size();
    }
    if ((rc = dupload(A, &A->am.uOff, uOff)) || (rc = dupload(A, &A->am.uCol, uCol)) ||
        (rc = dupload(A, &A->am.uPos, uPos)))
      return rc;
  }
  for (double** p : {&A->ax, &A->axo, &A->asol, &A->uA, &A->uB, &A->uC, &A->uc, &A->bbar, &A->aA, &A->rhs0bar,
                     &A->axn, &A->colR[0], &A->colR[1], &A->xbarA, &A->vbarA, &A->xbarB, &A->vbarB, &A->endX,
                     &A->endV})
    if ((rc = dalloc(A, p, V3))) return rc;
  for (double** p : {&A->Dbar, &A->DbarEl, &A->colD[0], &A->colD[1]})
    if ((rc = dalloc(A, p, (size_t)m.nSim))) return rc;
  if ((rc = dalloc(A, &A->tetAdj, 12 * (size_t)std::max(1, m.nTets)))) return rc;
  if ((rc = dalloc(A, &A->Gbar, (size_t)m.sellEntries))) return rc;
  if ((rc = dalloc(A, &A->R1bar, hs.r1Col.size()))) return rc;
  if ((rc = dalloc(A, &A->sbar, hs.cNbr.size()))) return rc;
  if ((rc = dalloc(A, &A->partials, (size_t)m.nPartials + 1))) return rc;
  if ((rc = dalloc(A, &A->gscale, 1))) return rc;
  if ((rc = dalloc(A, &A->colRacc, V3))) return rc;
  A->lay.V3 = V3;
  A->lay.nSim = (size_t)m.nSim;
  A->lay.LG = h->prm.lg_iters;
  A->lay.K = h->prm.cheb_iters;
  if ((rc = dalloc(A, &A->stage, A->lay.bytes() / sizeof(double)))) return rc;
  A->uS.assign(h->prm.cheb_iters + 2, nullptr);
  for (auto& u : A->uS)
    if ((rc = dalloc(A, &u, V3))) return rc;
  return PDIPC_OK;
}

}  // namespace

// Re-uploads the adjoint's copies of stiffness-dependent values after
// pdipc_set_cloth_stiffness has replaced the SELL 2-path products (h->sell.pA).
int adjoint_refresh_values(pdipc_handle* h) {
  AdjointState* A = h->adj;
  if (!A || A->midQ.empty()) return PDIPC_OK;
  const SellLayout& L = h->sell;
  std::vector<double> midA(A->midQ.size());
  for (size_t at = 0; at < midA.size(); at++) midA[at] = L.pA[A->midQ[at]];
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(A->am.midA, midA.data(), midA.size() * sizeof(double), hipMemcpyHostToDevice));
  return PDIPC_OK;
}

namespace {

// replays `g` (capturing `body` on first use) or, without graphs, runs `body` directly
template <class F>
int run_captured(pdipc_handle* h, hipGraphExec_t& g, F&& body) {
  if (!h->prm.use_graphs || h->m.timer) {  // kernel timing times each launch: no graph replay
    body();
    return check_launch(h);
  }
  if (!g) {
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    body();
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

// the recorded local/global loop of one outer iteration, into the stage
void enqueue_taped_lg(pdipc_handle* h, int withCol) {
  AdjointState* A = h->adj;
  DeviceMesh& m = h->m;
  const pdipc_params& p = h->prm;
  hipStream_t s = h->stream;
  const StageLayout& L = A->lay;
  double* b = A->stage;
  launch_copy(L.xold(b), m.Xold, (int64_t)L.V3, s);
  launch_copy(L.xstart(b), m.X, (int64_t)L.V3, s);
  launch_diag(m, withCol, s);
  launch_copy(L.D(b), m.sysDiag, m.nSim, s);
  launch_operators(m, p.omega, PDIPC_FP64, s);
  std::vector<double*> outs(p.cheb_iters + 1);
  for (int l = 0; l < p.lg_iters; l++) {
    launch_local(m, PDIPC_FP64, p.sigma_lo, p.sigma_hi, withCol, s);
    launch_copy(L.rhs(b, l), m.sysRhs, (int64_t)L.V3, s);
    for (int k = 0; k <= p.cheb_iters; k++) outs[k] = L.it(b, l, k);
    launch_solve_taped(m, p.omega, p.rho, p.cheb_iters, outs.data(), s);
  }
}

int slot(AdjointState* A, size_t n, double** out) { return A->tape.alloc(n * sizeof(double), (void**)out); }

// One FP64 step with every quantity the reverse pass needs recorded.  The
// kernel sequence is the step's own (advance in pdipc_api.hip) with the
// Chebyshev iterates written to the tape and the CCD recording per-record
// TOIs, so the recorded trajectory is bit-identical to the plain step's.
int taped_step(pdipc_handle* h, StepTape& st) {
  AdjointState* A = h->adj;
  DeviceMesh& m = h->m;
  const pdipc_params& p = h->prm;
  hipStream_t s = h->stream;
  size_t V3 = 3 * (size_t)m.ld;
  int rc;
  if ((rc = slot(A, V3, &st.x0)) || (rc = slot(A, V3, &st.v0))) return rc;
  launch_copy(st.x0, m.X, (int64_t)V3, s);
  launch_copy(st.v0, m.V, (int64_t)V3, s);
  launch_copy(m.X0, m.X, (int64_t)V3, s);
  launch_copy(m.Xold, m.X, (int64_t)V3, s);
  launch_predict(m, p.dt, s);
  bool activeCol = false;
  int outer = 0;
  for (; outer < p.max_outer; outer++) {
    OuterTape ot;
    ot.withCol = activeCol ? 1 : 0;
    if ((rc = run_captured(h, A->fwdGraph[ot.withCol], [&]() { enqueue_taped_lg(h, ot.withCol); }))) return rc;
    if ((rc = A->tape.alloc(A->lay.bytes(), (void**)&ot.blob))) return rc;
    HIPCHK(hipMemcpyAsync(ot.blob, A->stage, A->lay.bytes(), hipMemcpyDeviceToDevice, s));
    launch_direction(m, s);
    if (outer == 0) {
      if ((rc = broad_phase(h, m.Xold, m.Dir))) return rc;
      st.nRecs = h->lastRecs;
      if ((rc = A->tape.alloc((size_t)std::max(1, st.nRecs) * sizeof(int4), (void**)&st.recs))) return rc;
      if (st.nRecs > 0)
        HIPCHK(hipMemcpyAsync(st.recs, m.recs, (size_t)st.nRecs * sizeof(int4), hipMemcpyDeviceToDevice, s));
      if (st.nRecs > A->recToiCap) {
        if (A->recToi) {
          (void)hipFree(A->recToi);
          A->allocs.erase(std::find(A->allocs.begin(), A->allocs.end(), (void*)A->recToi));
        }
        A->recToiCap = st.nRecs + st.nRecs / 2 + 1024;
        if ((rc = dalloc(A, &A->recToi, (size_t)A->recToiCap))) return rc;
      }
    }
    HIPCHK(hipMemsetAsync(m.counters + 4, 0x7f, sizeof(int), s));  // (launch_direction reset the TOI minimum)
    launch_ccd_records(m, m.recs, h->lastRecs, m.Xold, m.X, p.ccd_filter, p.ccd_iters, p.ccd_step, p.ccd_gap,
                       p.touch_policy, A->recToi, m.counters + 4, s);
    launch_clamp_residual(m, p.toi_scale, s);
    if ((rc = collision(h))) return rc;  // reads the scalars back
    activeCol = true;
    ot.nAct = h->hCounters()[2];
    if ((rc = A->tape.alloc((size_t)std::max(1, ot.nAct) * sizeof(int), (void**)&ot.act))) return rc;
    if (ot.nAct > 0)
      HIPCHK(hipMemcpyAsync(ot.act, m.actRec, (size_t)ot.nAct * sizeof(int), hipMemcpyDeviceToDevice, s));
    double res = *h->hResidual();
    double tmin;
    unsigned long long bits = *h->hToi();
    memcpy(&tmin, &bits, 8);
    int am = h->hCounters()[4];
    ot.clamped = tmin < 1.0 ? 1 : 0;
    ot.toi = tmin < 1.0 ? tmin * p.toi_scale : 1.0;
    ot.argmin = (ot.clamped && am != 0x7f7f7f7f) ? am : -1;
    st.outers.push_back(std::move(ot));
    if (!std::isfinite(res)) return fail(PDIPC_E_NAN, "non-finite search direction");
    if (res < p.residual_tol && outer > 0) break;
    launch_copy(m.Xold, m.X, (int64_t)V3, s);
  }
  launch_velocity(m, p.dt, s);
  A->stats[0] += (double)st.outers.size();
  return check_launch(h);
}

// Chebyshev weights of the sweeps: W[0] = 1 (x1 = M^2 x0 + c), W[1] = 2/(2 - r),
// W[k+1] = 4/(4 - r W[k])
std::vector<double> cheb_weights(const pdipc_params& p) {
  std::vector<double> W(p.cheb_iters + 1, 0.0);
  double r = p.rho * p.rho;
  W[0] = 1.0;
  if (p.cheb_iters > 1) W[1] = 2.0 / (2.0 - r);
  for (int k = 1; k + 1 < p.cheb_iters; k++) W[k + 1] = 4.0 / (4.0 - r * W[k]);
  return W;
}

// Reverse of one recorded step: (xbarIn, vbarIn) = adjoints of the state after
// the step -> (xbarOut, vbarOut) = adjoints of the state before it; the
// stiffness adjoints accumulate in sbar, R1bar and DbarEl.
int reverse_step(pdipc_handle* h, StepTape& st, const double* xbarIn, const double* vbarIn, double* xbarOut,
                 double* vbarOut) {
  AdjointState* A = h->adj;
  DeviceMesh& m = h->m;
  const pdipc_params& p = h->prm;
  hipStream_t s = h->stream;
  size_t V3 = 3 * (size_t)m.ld;
  const std::vector<double> W = cheb_weights(p);
  const int K = p.cheb_iters, LG = p.lg_iters;
  double invt = 1.0 / p.dt * (1.0 - p.dt);
  // v = (x - x0) (1 - dt) / dt
  launch_axpby(m, 1.0, xbarIn, invt, vbarIn, A->ax, s);
  launch_axpby(m, -invt, vbarIn, 0.0, nullptr, A->axn, s);
  launch_zero(A->rhs0bar, (int64_t)V3, s);
  int cur = 0;
  launch_zero(A->colD[cur], m.nSim, s);
  launch_zero(A->colR[cur], (int64_t)V3, s);
  const int O = (int)st.outers.size();
  const StageLayout& L = A->lay;
  double* sg = A->stage;
  for (int o = O - 1; o >= 0; o--) {
    OuterTape& ot = st.outers[o];
    HIPCHK(hipMemcpyAsync(sg, ot.blob, L.bytes(), hipMemcpyDeviceToDevice, s));
    const double* xsol = L.it(sg, LG - 1, K);
    const double* xold = L.xold(sg);
    // collision terms computed at the end of this outer iteration (read by the next)
    if (o < O - 1)
      launch_adj_collision(m, st.recs, ot.act, ot.nAct, L.xold(st.outers[o + 1].blob), p.dhat, p.kappa, p.dt,
                           A->colD[cur], A->colR[cur], A->ax, s);
    // x = xold + toi (xsol - xold)
    launch_axpby(m, ot.toi, A->ax, 0.0, nullptr, A->asol, s);
    launch_axpby(m, 1.0 - ot.toi, A->ax, 0.0, nullptr, A->axo, s);
    if (ot.clamped && ot.argmin >= 0) {
      launch_dot_dir(m, A->ax, xsol, xold, p.toi_scale, A->partials, A->gscale, s);
      launch_adj_ccd(m, st.recs, ot.argmin, xold, xsol, p.ccd_filter, p.ccd_iters, p.ccd_step, p.ccd_gap,
                     p.touch_policy, A->gscale, A->axo, A->asol, s);
    }
    int nxt = 1 - cur;
    launch_zero(A->Gbar, m.sellEntries, s);
    launch_zero(A->Dbar, m.nSim, s);
    if (ot.withCol) launch_zero(A->colRacc, (int64_t)V3, s);
    // the operators of this outer iteration, from its recorded diagonal
    launch_copy(m.sysDiag, L.D(sg), m.nSim, s);
    HIPCHK(hipMemsetAsync(m.dChanged, 1, m.nSim, s));
    HIPCHK(hipMemsetAsync(m.dStale, 1, m.nSim, s));  // the next plain step recomputes every row
    launch_operators(m, p.omega, PDIPC_FP64, s);
    int rc = run_captured(h, A->revGraph[ot.withCol], [&]() {
      for (int l = LG - 1; l >= 0; l--) {
        launch_div_rows(m, A->asol, m.sysDiag, A->uS[K], s);  // u_K
        launch_zero(A->uc, (int64_t)V3, s);
        launch_adj_pass(m, A->am, W.data(), K, A->uS.data(), L.it(sg, l, 0), L.it(sg, l, 1), (int64_t)V3,
                        A->Gbar, A->uc, s);
        launch_adj_first(m, p.omega, L.rhs(sg, l), A->uc, A->bbar, A->Dbar, A->R1bar, s);
        launch_adj_local(m, A->am, p.sigma_lo, p.sigma_hi, L.it(sg, l, 0), A->bbar, A->uS[0], A->aA, A->tetAdj,
                         A->asol, A->sbar, A->rhs0bar, ot.withCol ? A->colRacc : nullptr, s);
      }
    });
    if (rc) return rc;
    launch_adj_ops_gather(m, A->am, p.omega, A->Gbar, A->Dbar, A->R1bar, s);
    launch_add_rows(A->DbarEl, A->Dbar, m.nSim, s);
    if (o > 0) {
      launch_copy(A->colD[nxt], A->Dbar, m.nSim, s);
      launch_copy(A->colR[nxt], A->colRacc, (int64_t)V3, s);
      launch_axpby(m, 1.0, A->axo, 1.0, A->asol, A->ax, s);  // both read x_clamp of outer o-1
      cur = nxt;
    } else {
      launch_axpby(m, 1.0, A->axn, 1.0, A->axo, A->uC, s);  // xold of the first outer iteration is x_n
      launch_predict_adj(m, p.dt, A->asol, A->rhs0bar, A->uC, xbarOut, vbarOut, s);
    }
    if ((rc = check_launch(h))) return rc;
  }
  return PDIPC_OK;
}

}  // namespace
}  // namespace pdipc

using namespace pdipc;

extern "C" {

int pdipc_rollout(pdipc_handle* h, int nsteps, int every) {
  int rc = require_device(h);
  if (rc) return rc;
  if (nsteps < 1 || every < 0) return fail(PDIPC_E_ARG, "nsteps must be positive, checkpoint_every >= 0");
  if (h->prm.cheb_iters > kMaxCheb) return fail(PDIPC_E_ARG, "the adjoint supports at most 32 Chebyshev sweeps");
  if (h->prm.precision != PDIPC_FP64) return fail(PDIPC_E_ARG, "the adjoint runs on FP64 steps");
  if ((rc = ensure_state(h))) return rc;
  AdjointState* A = h->adj;
  DeviceMesh& m = h->m;
  size_t V3 = 3 * (size_t)m.ld;
  if (every == 0) {  // record the whole run: the backward reverses this tape directly
    A->nSteps = nsteps;
    A->every = 0;
    A->recorded = false;
    A->lastKept = false;
    A->tape.reset();
    A->stats[0] = 0;
    A->seg.assign(nsteps, StepTape());
    for (int n = 0; n < nsteps; n++)
      if ((rc = taped_step(h, A->seg[n]))) return rc;
    double used = 0;
    for (size_t c = 0; c < A->tape.ci && c < A->tape.chunks.size(); c++) used += (double)A->tape.chunks[c].size;
    used += (double)A->tape.off;
    A->stats[1] = used;
    A->recorded = true;
    launch_copy(A->endX, m.X, (int64_t)V3, h->stream);
    launch_copy(A->endV, m.V, (int64_t)V3, h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
    return check_launch(h);
  }
  A->recorded = false;
  A->lastKept = false;
  int nck = (nsteps + every - 1) / every;
  while ((int)A->ckX.size() < nck) {
    double *x = nullptr, *v = nullptr;
    HIPCHK(hipMalloc((void**)&x, V3 * sizeof(double)));
    A->ckX.push_back(x);
    HIPCHK(hipMalloc((void**)&v, V3 * sizeof(double)));
    A->ckV.push_back(v);
  }
  A->nSteps = nsteps;
  A->every = every;
  // the last segment runs recorded (the same trajectory bit for bit), so the
  // backward reverses it without re-running it
  const int keepFrom = (nck - 1) * every;
  A->tape.reset();
  A->seg.assign(nsteps - keepFrom, StepTape());
  for (int n = 0; n < nsteps; n++) {
    if (n % every == 0) {
      launch_copy(A->ckX[n / every], m.X, (int64_t)V3, h->stream);
      launch_copy(A->ckV[n / every], m.V, (int64_t)V3, h->stream);
    }
    if (n >= keepFrom)
      rc = taped_step(h, A->seg[n - keepFrom]);
    else
      rc = advance(h);
    if (rc) return rc;
  }
  A->lastKept = true;
  launch_copy(A->endX, m.X, (int64_t)V3, h->stream);
  launch_copy(A->endV, m.V, (int64_t)V3, h->stream);
  HIPCHK(hipStreamSynchronize(h->stream));
  return check_launch(h);
}

int pdipc_backward(pdipc_handle* h, const double* dx, const double* dv, double* dx0, double* dv0, double* dk, int nk) {
  int rc = require_device(h);
  if (rc) return rc;
  AdjointState* A = h->adj;
  if (!A || A->nSteps < 1) return fail(PDIPC_E_ARG, "pdipc_backward needs a pdipc_rollout first");
  if (h->prm.precision != PDIPC_FP64) return fail(PDIPC_E_ARG, "the adjoint runs on FP64 steps");
  if (dk && nk < h->hs.nClothTris) return fail(PDIPC_E_ARG, "dL/dk holds fewer entries than the cloth triangles");
  DeviceMesh& m = h->m;
  const HostSetup& hs = h->hs;
  hipStream_t s = h->stream;
  size_t V3 = 3 * (size_t)m.ld;
  // seeds (AoS over the simulated points; null = zero)
  std::vector<double> zero(3 * (size_t)m.nSim, 0.0);
  if ((rc = put_soa(h, A->xbarA, dx ? dx : zero.data(), m.nSim))) return rc;
  if ((rc = put_soa(h, A->vbarA, dv ? dv : zero.data(), m.nSim))) return rc;
  launch_zero(A->sbar, (int64_t)hs.cNbr.size(), s);
  launch_zero(A->R1bar, (int64_t)hs.r1Col.size(), s);
  launch_zero(A->DbarEl, m.nSim, s);
  double *xin = A->xbarA, *vin = A->vbarA, *xout = A->xbarB, *vout = A->vbarB;
  A->stats[2] = 0;
  if (A->recorded) {  // the rollout's own tape, every step
    for (int n = A->nSteps - 1; n >= 0; n--) {
      if ((rc = reverse_step(h, A->seg[n], xin, vin, xout, vout))) return rc;
      std::swap(xin, xout);
      std::swap(vin, vout);
    }
  }
  A->stats[0] = A->recorded ? A->stats[0] : 0;
  double tapeBytes = A->recorded ? A->stats[1] : 0;
  int nseg = A->recorded ? 0 : (A->nSteps + A->every - 1) / A->every;
  for (int sg = nseg - 1; sg >= 0; sg--) {
    int first = sg * A->every, last = std::min(A->nSteps, first + A->every);
    if (A->lastKept && sg == nseg - 1) {
      for (const StepTape& st : A->seg) A->stats[0] += (double)st.outers.size();
    } else {
      launch_copy(m.X, A->ckX[sg], (int64_t)V3, s);
      launch_copy(m.V, A->ckV[sg], (int64_t)V3, s);
      A->lastKept = false;  // the tape now holds this segment
      A->tape.reset();
      A->seg.assign(last - first, StepTape());
      for (int n = first; n < last; n++)
        if ((rc = taped_step(h, A->seg[n - first]))) return rc;
    }
    double used = 0;
    for (size_t c = 0; c < A->tape.ci && c < A->tape.chunks.size(); c++) used += (double)A->tape.chunks[c].size;
    used += (double)A->tape.off;
    tapeBytes = std::max(tapeBytes, used);
    for (int n = last - 1; n >= first; n--) {
      if ((rc = reverse_step(h, A->seg[n - first], xin, vin, xout, vout))) return rc;
      std::swap(xin, xout);
      std::swap(vin, vout);
    }
    A->stats[2] += 1;
  }
  A->stats[1] = tapeBytes;
  A->stats[3] = A->nSteps;
  HIPCHK(hipStreamSynchronize(s));
  if ((rc = check_launch(h))) return rc;
  if (dx0 && (rc = get_soa(h, xin, dx0, m.nSim))) return rc;
  if (dv0 && (rc = get_soa(h, vin, dv0, m.nSim))) return rc;
  if (dk) {
    std::vector<double> sb(hs.cNbr.size()), rb(hs.r1Col.size()), de(m.nSim);
    HIPCHK(hipMemcpy(sb.data(), A->sbar, sb.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(rb.data(), A->R1bar, rb.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(de.data(), A->DbarEl, de.size() * sizeof(double), hipMemcpyDeviceToHost));
    // spring weight of a slot enters: the local step (sbar), R1(i, nbr) and the diagonal D_i
    std::vector<double> eg(hs.edgeTriOff.size(), 0.0);
    for (size_t c = 0; c < hs.cCenter.size(); c++) {
      int i = hs.cCenter[c];
      for (int sl = hs.cOff[c] + 1; sl < hs.cOff[c + 1]; sl++) {
        double g = sb[sl] + de[i] + (hs.cSlotR1[sl] >= 0 ? rb[hs.cSlotR1[sl]] : 0.0);
        eg[hs.cSlotEdge[sl]] += g;
      }
    }
    std::fill(dk, dk + hs.nClothTris, 0.0);
    for (size_t e = 0; e < eg.size(); e++) {
      int o = hs.edgeTriOff[e];
      int n = (e + 1 < eg.size() ? hs.edgeTriOff[e + 1] : (int)hs.edgeTriList.size()) - o;
      for (int q = 0; q < n; q++) dk[hs.edgeTriList[o + q]] += eg[e] / n;
    }
  }
  // leave the handle at the end of the rollout
  launch_copy(m.X, A->endX, (int64_t)V3, s);
  launch_copy(m.V, A->endV, (int64_t)V3, s);
  HIPCHK(hipStreamSynchronize(s));
  return check_launch(h);
}

int pdipc_debug_solve_vjp(pdipc_handle* h, const double* diag, const double* rhs, const double* x0,
                          const double* xbar_out, double* xbar0, double* bbar) {
  int rc = require_device(h);
  if (rc) return rc;
  if ((rc = ensure_state(h))) return rc;
  AdjointState* A = h->adj;
  DeviceMesh& m = h->m;
  const pdipc_params& p = h->prm;
  hipStream_t s = h->stream;
  size_t V3 = 3 * (size_t)m.ld;
  HIPCHK(hipMemcpyAsync(m.sysDiag, diag, m.nSim * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemsetAsync(m.dChanged, 1, m.nSim, s));
  HIPCHK(hipMemsetAsync(m.dStale, 1, m.nSim, s));
  if ((rc = put_soa(h, A->bbar, rhs, m.nSim))) return rc;
  if ((rc = put_soa(h, A->axo, x0, m.nTotal))) return rc;
  if ((rc = put_soa(h, A->asol, xbar_out, m.nSim))) return rc;
  launch_operators(m, p.omega, PDIPC_FP64, s);
  const std::vector<double> W = cheb_weights(p);
  const int K = p.cheb_iters;
  launch_div_rows(m, A->asol, m.sysDiag, A->uS[K], s);
  launch_zero(A->uc, (int64_t)V3, s);
  launch_zero(A->Gbar, m.sellEntries, s);
  launch_adj_pass(m, A->am, W.data(), K, A->uS.data(), A->axo, A->axo, 0, A->Gbar, A->uc, s);
  launch_zero(A->Dbar, m.nSim, s);
  launch_adj_first(m, p.omega, A->bbar, A->uc, A->ax, A->Dbar, A->R1bar, s);
  launch_mul_rows(m, A->uS[0], m.sysDiag, A->asol, s);
  if ((rc = check_launch(h))) return rc;
  if ((rc = get_soa(h, A->asol, xbar0, m.nSim))) return rc;
  return get_soa(h, A->ax, bbar, m.nSim);
}

int pdipc_debug_local_vjp(pdipc_handle* h, const double* x, const double* bbar, double* xbar) {
  int rc = require_device(h);
  if (rc) return rc;
  if ((rc = ensure_state(h))) return rc;
  AdjointState* A = h->adj;
  DeviceMesh& m = h->m;
  hipStream_t s = h->stream;
  size_t V3 = 3 * (size_t)m.ld;
  if ((rc = put_soa(h, A->axo, x, m.nTotal))) return rc;
  if ((rc = put_soa(h, A->bbar, bbar, m.nSim))) return rc;
  launch_zero(A->uB, (int64_t)V3, s);
  launch_zero(A->rhs0bar, (int64_t)V3, s);
  launch_adj_local(m, A->am, h->prm.sigma_lo, h->prm.sigma_hi, A->axo, A->bbar, A->uB, A->aA, A->tetAdj, A->asol,
                   A->sbar, A->rhs0bar, nullptr, s);
  if ((rc = check_launch(h))) return rc;
  return get_soa(h, A->asol, xbar, m.nSim);
}

int pdipc_debug_tape(pdipc_handle* h, int step, int outer, int which, int l, int k, double* out) {
  int rc = require_device(h);
  if (rc) return rc;
  AdjointState* A = h->adj;
  if (!A || step < 0 || step >= (int)A->seg.size()) return fail(PDIPC_E_ARG, "no such recorded step");
  StepTape& st = A->seg[step];
  if (outer < 0 || outer >= (int)st.outers.size()) return fail(PDIPC_E_ARG, "no such outer iteration");
  OuterTape& ot = st.outers[outer];
  if (which == 5) {  // scalars: toi, clamped, argmin, collision terms active
    out[0] = ot.toi;
    out[1] = ot.clamped;
    out[2] = ot.argmin;
    out[3] = ot.withCol;
    return PDIPC_OK;
  }
  const StageLayout& L = A->lay;
  const double* src = nullptr;
  switch (which) {
    case 0: src = L.xold(ot.blob); break;
    case 1: src = L.xstart(ot.blob); break;
    case 2: src = L.D(ot.blob); break;
    case 3: src = (l >= 0 && l < L.LG) ? L.rhs(ot.blob, l) : nullptr; break;
    default: src = (l >= 0 && l < L.LG && k >= 0 && k <= L.K) ? L.it(ot.blob, l, k) : nullptr;
  }
  if (!src) return fail(PDIPC_E_ARG, "no such recorded vector");
  if (which == 2) {
    HIPCHK(hipMemcpy(out, src, h->m.nSim * sizeof(double), hipMemcpyDeviceToHost));
    return PDIPC_OK;
  }
  return get_soa(h, src, out, h->m.nSim);
}

int pdipc_adjoint_stats(pdipc_handle* h, double* out4) {
  if (!h || !out4) return fail(PDIPC_E_ARG, "null argument");
  for (int i = 0; i < 4; i++) out4[i] = h->adj ? h->adj->stats[i] : 0.0;
  return PDIPC_OK;
}

}  // extern "C"
