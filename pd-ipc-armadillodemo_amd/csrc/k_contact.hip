// Narrow phase over the candidate records, all in FP64:
//   k_ccd        ACCD time of impact per record (ccdVertexTriangleToi /
//                ccdEdgeEdgeToi, Simulator/PD-IPC/PdIpcGpuFunc.cu:2460-2607) behind the
//                swept-box filter (:2346-2458), reduced to one device-resident
//                minimum (wavefront shuffle min + one 64-bit atomicMin per wave;
//                non-negative doubles order like their bit patterns).
//   k_collision  the PD barrier projection (setCollisionConstraintsKernel,
//                :2756-2892): weight = barrier(d^2) + min(m_p, m_q)/dt^2 on the
//                four diagonal entries, weight * target on the four rhs rows.
// Records are int4: point-triangle (p, t0, t1, t2); edge-edge (~a0, a1, b0, b1).
#include "device.hpp"
#include "dual.hpp"

namespace pdipc {
namespace {

constexpr int kBlock = 256;

template <class T>
struct P3T { T v[3]; };
using P3 = P3T<double>;

template <class T>
__host__ __device__ inline P3T<T> sub(const P3T<T>& a, const P3T<T>& b) { return {{a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]}}; }
template <class T>
__host__ __device__ inline T dot(const P3T<T>& a, const P3T<T>& b) { return a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2]; }
template <class T>
__host__ __device__ inline P3T<T> cross(const P3T<T>& a, const P3T<T>& b) {
  return {{a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2], a.v[0] * b.v[1] - a.v[1] * b.v[0]}};
}

// The primitive distances, ACCD and barrier projection are templated on the
// scalar: double in the step, Dual<N> in the adjoint (same branches, exact
// derivatives of the computed function).

// VertexEdgeSqDistance (PdIpcGpuFunc.cu:2155-2174)
template <class T>
__host__ __device__ inline T ve_d2(const P3T<T>& v, const P3T<T>& ea, const P3T<T>& eb, T* r) {
  P3T<T> va = sub(v, ea), ba = sub(eb, ea);
  T vb = dot(va, ba), bb = dot(ba, ba);
  *r = vb < T(0.0) ? T(0.0) : (vb > bb ? T(1.0) : vb / bb);
  P3T<T> d;
#pragma unroll
  for (int k = 0; k < 3; k++) d.v[k] = v.v[k] - (ea.v[k] * (T(1.0) - *r) + eb.v[k] * *r);
  return dot(d, d);
}

// EdgeEdgeSqDistance (PdIpcGpuFunc.cu:2176-2236); parallel edges (cross product
// below 1e-24) take the minimum over all four endpoint-edge distances.
template <class T>
__host__ __device__ inline T ee_d2(const P3T<T>& xi, const P3T<T>& xj, const P3T<T>& xa, const P3T<T>& xb, T* r, T* s,
                          P3T<T>* dir) {
  P3T<T> xba = sub(xb, xa), xji = sub(xj, xi), xai = sub(xa, xi);
  P3T<T> n = cross(xji, xba);
  T nn = dot(n, n);
  T waiji = dot(n, cross(xai, xji));
  T waiba = dot(n, cross(xai, xba));
  if (nn > T(1e-24) && waiji >= T(0.0) && waiji <= nn && waiba >= T(0.0) && waiba <= nn) {
    *r = waiba / nn;
    *s = waiji / nn;
  } else {
    bool dg = !(nn > T(1e-24));
    T best = T(999999999.0), d, v;
    *r = T(0.0); *s = T(0.0);
    if ((dg || waiba < T(0.0)) && (d = ve_d2(xi, xa, xb, &v)) < best) { best = d; *r = T(0.0); *s = v; }
    if ((dg || waiba > nn) && (d = ve_d2(xj, xa, xb, &v)) < best) { best = d; *r = T(1.0); *s = v; }
    if ((dg || waiji < T(0.0)) && (d = ve_d2(xa, xi, xj, &v)) < best) { best = d; *r = v; *s = T(0.0); }
    if ((dg || waiji > nn) && (d = ve_d2(xb, xi, xj, &v)) < best) { best = d; *r = v; *s = T(1.0); }
  }
#pragma unroll
  for (int k = 0; k < 3; k++)
    dir->v[k] = xi.v[k] * (T(1.0) - *r) + xj.v[k] * *r - xa.v[k] * (T(1.0) - *s) - xb.v[k] * *s;
  return dot(*dir, *dir);
}

// VertexTriangleDistance (PdIpcGpuFunc.cu:2238-2292); degenerate triangles take
// the minimum over the three edges.
template <class T>
__host__ __device__ inline T pt_d2(const P3T<T>& xi, const P3T<T>& xa, const P3T<T>& xb, const P3T<T>& xc, T* ba, T* bb, T* bc,
                          P3T<T>* dir) {
  P3T<T> xba = sub(xb, xa), xca = sub(xc, xa), xia = sub(xi, xa);
  P3T<T> n = cross(xba, xca);
  T nn = dot(n, n);
  T wiaca = dot(n, cross(xia, xca));
  T wbaia = dot(n, cross(xba, xia));
  if (nn > T(1e-24) && wiaca >= T(0.0) && wbaia >= T(0.0) && nn - wiaca - wbaia >= T(0.0)) {
    *bb = wiaca / nn;
    *bc = wbaia / nn;
    *ba = T(1.0) - *bb - *bc;
  } else {
    bool dg = !(nn > T(1e-24));
    T best = T(999999999.0), r, d;
    *ba = T(1.0); *bb = T(0.0); *bc = T(0.0);
    if ((dg || nn - wiaca - wbaia < T(0.0)) && (d = ve_d2(xi, xb, xc, &r)) < best) {
      best = d; *bb = T(1.0) - r; *bc = r; *ba = T(0.0);
    }
    if ((dg || wiaca < T(0.0)) && (d = ve_d2(xi, xa, xc, &r)) < best) {
      best = d; *bb = T(0.0); *bc = r; *ba = T(1.0) - *bb - *bc;
    }
    if ((dg || wbaia < T(0.0)) && (d = ve_d2(xi, xa, xb, &r)) < best) {
      best = d; *bb = r; *bc = T(0.0); *ba = T(1.0) - *bb - *bc;
    }
  }
#pragma unroll
  for (int k = 0; k < 3; k++) dir->v[k] = xi.v[k] - xa.v[k] * *ba - xb.v[k] * *bb - xc.v[k] * *bc;
  return dot(*dir, *dir);
}

template <class T> __host__ __device__ inline T mn2(T a, T b) { return a < b ? a : b; }
template <class T> __host__ __device__ inline T mx2(T a, T b) { return a > b ? a : b; }
// Bounds of the swept-box and position filters as fmin / fmax (one FP64
// instruction each instead of a compare and two selects): for finite
// coordinates they pick the same values as the reference's ternary min2 /
// max2 (DataCore.h:32), so every filter decision is the same.
__host__ __device__ inline double bmn(double a, double b) { return fmin(a, b); }
__host__ __device__ inline double bmx(double a, double b) { return fmax(a, b); }
__host__ __device__ inline double bmn3(double a, double b, double c) { return fmin(fmin(a, b), c); }
__host__ __device__ inline double bmx3(double a, double b, double c) { return fmax(fmax(a, b), c); }

__host__ __device__ inline bool pt_ccd_filter(const P3* x, const P3* dx, double dist) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double tmin = bmn3(x[1].v[k], x[2].v[k], x[3].v[k]);
    tmin = bmn3(tmin, x[1].v[k] + dx[1].v[k], x[2].v[k] + dx[2].v[k]);
    tmin = bmn(tmin, x[3].v[k] + dx[3].v[k]);
    double tmax = bmx3(x[1].v[k], x[2].v[k], x[3].v[k]);
    tmax = bmx3(tmax, x[1].v[k] + dx[1].v[k], x[2].v[k] + dx[2].v[k]);
    tmax = bmx(tmax, x[3].v[k] + dx[3].v[k]);
    double p0 = x[0].v[k], p1 = x[0].v[k] + dx[0].v[k];
    double pmin = bmn(p0, p1), pmax = bmx(p0, p1);
    if ((pmin - tmax) > dist || (tmin - pmax) > dist) return false;
  }
  return true;
}
__host__ __device__ inline bool ee_ccd_filter(const P3* x, const P3* dx, double dist) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double mina = bmn(x[0].v[k], x[0].v[k] + dx[0].v[k]);
    mina = bmn(mina, x[1].v[k]);
    mina = bmn(mina, x[1].v[k] + dx[1].v[k]);
    double maxa = bmx(x[0].v[k], x[0].v[k] + dx[0].v[k]);
    maxa = bmx(maxa, x[1].v[k]);
    maxa = bmx(maxa, x[1].v[k] + dx[1].v[k]);
    double minb = bmn(x[2].v[k], x[2].v[k] + dx[2].v[k]);
    minb = bmn(minb, x[3].v[k]);
    minb = bmn(minb, x[3].v[k] + dx[3].v[k]);
    double maxb = bmx(x[2].v[k], x[2].v[k] + dx[2].v[k]);
    maxb = bmx(maxb, x[3].v[k]);
    maxb = bmx(maxb, x[3].v[k] + dx[3].v[k]);
    if ((mina - maxb) > dist || (minb - maxa) > dist) return false;
  }
  return true;
}

template <class T>
__host__ __device__ inline T ccd_dfun(int type, const P3T<T>* x, T* d2out) {
  T a, b, c;
  P3T<T> dir;
  if (type == 0) {
    T d2 = pt_d2(x[0], x[1], x[2], x[3], &a, &b, &c, &dir);
    *d2out = d2;
    return d2;
  }
  T d2 = ee_d2(x[0], x[1], x[2], x[3], &a, &b, &dir);
  if (d2 <= T(0.0)) {
    P3T<T> e = sub(x[0], x[2]);
    d2 = dot(e, e);
    e = sub(x[0], x[3]); d2 = mn2(dot(e, e), d2);
    e = sub(x[1], x[2]); d2 = mn2(dot(e, e), d2);
    e = sub(x[1], x[3]); d2 = mn2(dot(e, e), d2);
  }
  *d2out = d2;
  return d2;
}

// ACCD with the reference's constants as parameters.  With `exitPath`, also
// the discrete path the loop took: 8 * (loop iteration) + how it ended (0 all
// iterations, 1 the gap test, 2 the clamp at 1, 3 a touch inside the loop,
// 4 no relative motion, 5 touching at the start, 6 the first-step return).
template <class T>
__host__ __device__ inline T accd(int type, P3T<T>* x, P3T<T>* dx, int iters, double step, double gapf, int touchPolicy,
                                  int* exitPath = nullptr) {
  P3T<T> mov;
#pragma unroll
  for (int k = 0; k < 3; k++) mov.v[k] = (dx[0].v[k] + dx[1].v[k] + dx[2].v[k] + dx[3].v[k]) / T(4.0);
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) dx[i].v[k] -= mov.v[k];
  T maxdisp;
  if (type == 0) {
    T m2 = mx2(dot(dx[1], dx[1]), dot(dx[2], dx[2]));
    m2 = mx2(m2, dot(dx[3], dx[3]));
    maxdisp = sqrt(dot(dx[0], dx[0])) + sqrt(m2);
  } else {
    T m0 = mx2(dot(dx[0], dx[0]), dot(dx[1], dx[1]));
    T m1 = mx2(dot(dx[2], dx[2]), dot(dx[3], dx[3]));
    maxdisp = sqrt(m0) + sqrt(m1);
  }
  if (maxdisp == T(0.0)) {
    if (exitPath) *exitPath = 4;
    return T(1.0);
  }
  T d2;
  T dF = ccd_dfun(type, x, &d2);
  T dcur = sqrt(d2);
  if (touchPolicy == 1 && !(dcur > T(0.0))) {
    if (exitPath) *exitPath = 5;
    return T(0.0);
  }
  T gap = T(gapf) * dF / dcur;
  T toc = T(0.0);
  if (exitPath) *exitPath = 8 * iters;
  for (int i = 0; i < iters; i++) {
    T tlb = T(step) * dF / (dcur * maxdisp);
    // First iteration with tlb > 1: the loop returns 1.  The move shifts the
    // primitives relative to each other by at most tlb * maxdisp = step * d0
    // (dF = d^2 for both record types), so the moved distance is at least
    // (1 - step) d0 > 0 and the touch test cannot fire; for d0 above 1e-9 no
    // rounding can turn that distance into 0.  The second distance
    // evaluation is skipped (most records are far apart relative to their
    // motion).
    if (i == 0 && tlb > T(1.0) && dcur > T(1e-9)) {
      if (exitPath) *exitPath = 6;
      return T(1.0);
    }
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int k = 0; k < 3; k++) x[j].v[k] += tlb * dx[j].v[k];
    dF = ccd_dfun(type, x, &d2);
    dcur = sqrt(d2);
    if (touchPolicy == 1 && !(dcur > T(0.0))) {
      if (exitPath) *exitPath = 8 * i + 3;
      return toc;
    }
    if (toc != T(0.0) && (dF / dcur) < gap) {
      if (exitPath) *exitPath = 8 * i + 1;
      break;
    }
    toc += tlb;
    if (toc > T(1.0)) {
      if (exitPath) *exitPath = 8 * i + 2;
      toc = T(1.0);
      break;
    }
  }
  return toc;
}

__host__ __device__ inline P3 load3(const double* X, int ld, int i) { return {{X[i], X[ld + i], X[2 * ld + i]}}; }

__global__ void k_toi_init(unsigned long long* bits) { *bits = (unsigned long long)__double_as_longlong(1.0); }

// One record's ACCD time of impact between positions X0 and X1 (1 when the
// record shares a vertex or fails the swept-box filter).  Component c of point
// i is X[c * cs + i * ps]: cs = ld, ps = 1 for the SoA state vectors, cs = 1,
// ps = 3 for AoS arrays.
__device__ __forceinline__ double record_toi(int4 r, int cs, int ps, const double* __restrict__ X0,
                                             const double* __restrict__ X1, double filter, int iters, double step,
                                             double gapf, int touchPolicy) {
  int type = r.x < 0 ? 1 : 0;
  int id[4] = {type ? ~r.x : r.x, r.y, r.z, r.w};
  bool skip = type == 0 ? (id[0] == id[1] || id[0] == id[2] || id[0] == id[3])
                        : (id[0] == id[2] || id[0] == id[3] || id[1] == id[2] || id[1] == id[3]);
  if (skip) return 1.0;
  P3 x[4], dx[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int64_t b = (int64_t)id[q] * ps;
    x[q] = {{X0[b], X0[b + cs], X0[b + 2 * (int64_t)cs]}};
    P3 y = {{X1[b], X1[b + cs], X1[b + 2 * (int64_t)cs]}};
    dx[q] = sub(y, x[q]);
  }
  bool pass = type == 0 ? pt_ccd_filter(x, dx, filter) : ee_ccd_filter(x, dx, filter);
  return pass ? accd(type, x, dx, iters, step, gapf, touchPolicy) : 1.0;
}

__device__ __forceinline__ double wave_min(double toi) {
  for (int o = kWave / 2; o > 0; o >>= 1) {
    double other = __shfl_xor(toi, o, kWave);
    toi = other < toi ? other : toi;
  }
  return toi;
}

__global__ void __launch_bounds__(kBlock) k_ccd(int nRecs, int ld, const int4* __restrict__ recs,
                                                const double* __restrict__ X0, const double* __restrict__ X1,
                                                double filter, int iters, double step, double gapf, int touchPolicy,
                                                unsigned long long* __restrict__ toiBits) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  double toi = 1.0;
  if (t < nRecs) {
    double v = record_toi(recs[t], ld, 1, X0, X1, filter, iters, step, gapf, touchPolicy);
    if (v < toi) toi = v;  // NaN (reference rule) never wins
  }
  toi = wave_min(toi);
  if ((threadIdx.x % kWave) == 0 && toi < 1.0) atomicMin(toiBits, (unsigned long long)__double_as_longlong(toi));
}

// The same minimum without atomics on AoS positions: each workgroup writes its
// own minimum to blockMin[blockIdx.x] (every entry of the grid is written).
__global__ void __launch_bounds__(kBlock) k_ccd_blockmin(int nRecs, const int4* __restrict__ recs,
                                                         const double* __restrict__ X0, const double* __restrict__ X1,
                                                         double filter, int iters, double step, double gapf,
                                                         int touchPolicy, double* __restrict__ blockMin) {
  __shared__ double waveMin[kBlock / kWave];
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  double toi = 1.0;
  if (t < nRecs) {
    double v = record_toi(recs[t], 1, 3, X0, X1, filter, iters, step, gapf, touchPolicy);
    if (v < toi) toi = v;
  }
  toi = wave_min(toi);
  if ((threadIdx.x % kWave) == 0) waveMin[threadIdx.x / kWave] = toi;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = waveMin[0];
    for (int w = 1; w < kBlock / kWave; w++) b = waveMin[w] < b ? waveMin[w] : b;
    blockMin[blockIdx.x] = b;
  }
}

// barrier (PdIpcGpuFunc.cu:2745-2754)
template <class T>
__host__ __device__ inline T barrier(T d2, double dh2, double kappa) {
  T e = T(0.0);
  if (d2 < T(dh2)) {
    T t2 = d2 - T(dh2);
    e = T(-kappa) * t2 * t2 * log(d2 / T(dh2)) / T(dh2 * dh2);
  }
  return e;
}

// the position filter of setCollisionConstraintsKernel (PdIpcGpuFunc.cu:2327-2423)
__host__ __device__ inline bool cd_filter(int type, const P3* x, double dHat) {
  bool pass = true;
  if (type == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double tmin = bmn3(x[1].v[k], x[2].v[k], x[3].v[k]), tmax = bmx3(x[1].v[k], x[2].v[k], x[3].v[k]);
      if ((x[0].v[k] - tmax) > dHat || (tmin - x[0].v[k]) > dHat) pass = false;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double mina = bmn(x[0].v[k], x[1].v[k]), maxa = bmx(x[0].v[k], x[1].v[k]);
      double minb = bmn(x[2].v[k], x[3].v[k]), maxb = bmx(x[2].v[k], x[3].v[k]);
      if ((mina - maxb) > dHat || (minb - maxa) > dHat) pass = false;
    }
  }
  return pass;
}

// One record's barrier projection (setCollisionConstraintsKernel,
// PdIpcGpuFunc.cu:2756-2892): returns false if inactive (d^2 >= dHat^2), else
// the weight w and the four targets x_q + shift_q * dir.
template <class T>
__host__ __device__ inline bool pair_terms(int type, const P3T<T>* x, const double* cmq, double dHat, double kappa, double dt,
                                  T* wOut, P3T<T>* tgt) {
  const double dh2 = dHat * dHat;
  P3T<T> dir;
  T d2, mp, mq;
  double sgn[4];
  if (type == 0) {
    T a, b, c;
    d2 = pt_d2(x[0], x[1], x[2], x[3], &a, &b, &c, &dir);
    mp = T(cmq[0]);
    mq = a * T(cmq[1]) + b * T(cmq[2]) + c * T(cmq[3]);
    sgn[0] = 1; sgn[1] = -1; sgn[2] = -1; sgn[3] = -1;
  } else {
    T rr, ss;
    d2 = ee_d2(x[0], x[1], x[2], x[3], &rr, &ss, &dir);
    mp = (T(1.0) - rr) * T(cmq[0]) + rr * T(cmq[1]);
    mq = (T(1.0) - ss) * T(cmq[2]) + ss * T(cmq[3]);
    sgn[0] = 1; sgn[1] = 1; sgn[2] = -1; sgn[3] = -1;
  }
  if (!(d2 < T(dh2))) return false;
  T dist = sqrt(d2);
  T w = barrier(d2, dh2, kappa) + mn2(mp, mq) / T(dt * dt);
  T rp = mq / (mp + mq), rq = mp / (mp + mq);
  T len = sqrt(dot(dir, dir));
#pragma unroll
  for (int k = 0; k < 3; k++) dir.v[k] /= len;
  T gapd = T(dHat) - dist;
  T dp = gapd * rp, dq = gapd * rq;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    T sh = sgn[q] > 0 ? dp : -dq;
#pragma unroll
    for (int k = 0; k < 3; k++) tgt[q].v[k] = x[q].v[k] + sh * dir.v[k];
  }
  *wOut = w;
  return true;
}

__host__ __device__ inline bool record_ids(int4 r, int id[4]) {  // false for shared-vertex pairs (skipped)
  int type = r.x < 0 ? 1 : 0;
  id[0] = type ? ~r.x : r.x; id[1] = r.y; id[2] = r.z; id[3] = r.w;
  bool skip = type == 0 ? (id[0] == id[1] || id[0] == id[2] || id[0] == id[3])
                        : (id[0] == id[2] || id[0] == id[3] || id[1] == id[2] || id[1] == id[3]);
  return !skip;
}

// The collision terms are summed per point in a fixed order (ascending record,
// then corner), so the step is bit-reproducible: no floating-point atomics.
//   1. k_col_terms: each active record appends (wave-aggregated) its weight and
//      four targets to a compact list and counts its simulated points;
//   2. exclusive scan of the per-point counts;
//   3. k_col_fill: every (point, record, corner) entry goes to its point's range;
//   4. k_col_rank: each entry moves to its rank by (record, corner) in the range;
//   5. k_col_gather: per point, the ordered range is summed.
__global__ void __launch_bounds__(kBlock) k_col_terms(int nRecs, int ld, int nSim, const int4* __restrict__ recs,
                                                      const double* __restrict__ X, const double* __restrict__ cm,
                                                      double dHat, double kappa, double dt, int actCap,
                                                      int* __restrict__ actRec, double* __restrict__ actW,
                                                      double* __restrict__ actT, int* __restrict__ ptCount,
                                                      int* __restrict__ counters) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int lane = threadIdx.x % kWave;
  bool active = false;
  double w = 0.0;
  P3 tgt[4];
  int id[4] = {0, 0, 0, 0};
  if (t < nRecs) {
    int4 r = recs[t];
    int type = r.x < 0 ? 1 : 0;
    if (record_ids(r, id)) {
      P3 x[4];
#pragma unroll
      for (int q = 0; q < 4; q++) x[q] = load3(X, ld, id[q]);
      if (cd_filter(type, x, dHat)) {
        double cmq[4] = {cm[id[0]], cm[id[1]], cm[id[2]], cm[id[3]]};
        active = pair_terms(type, x, cmq, dHat, kappa, dt, &w, tgt);
      }
    }
  }
  unsigned long long b = __ballot(active);
  if (!b) return;
  int leader = __ffsll((long long)b) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(counters + 2, (int)__popcll(b));
  base = __shfl(base, leader, kWave);
  if (!active) return;
  const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  int a = base + (int)__popcll(b & lt);
  if (a >= actCap) {
    atomicOr(counters + 5, 1);  // the host grows the list and repeats the pass
    return;
  }
  actRec[a] = t;
  actW[a] = w;
#pragma unroll
  for (int q = 0; q < 4; q++) {
#pragma unroll
    for (int k = 0; k < 3; k++) actT[(int64_t)a * 12 + 3 * q + k] = tgt[q].v[k];
    if (id[q] < nSim) atomicAdd(ptCount + id[q], 1);
  }
}

// zeroes the pass's counters (counters[2]: active records, counters[5]: list
// overflow) and the per-point counts and cursors: one launch instead of four
// memsets
__global__ void k_col_clear(int nSim, int* __restrict__ counters, int* __restrict__ ptCount,
                            int* __restrict__ ptCursor) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    counters[2] = 0;
    counters[5] = 0;
  }
  if (i <= nSim) ptCount[i] = 0;
  if (i < nSim) ptCursor[i] = 0;
}

__global__ void k_col_fill(int ld, int nSim, int actCap, const int4* __restrict__ recs, const int* __restrict__ counters,
                           const int* __restrict__ actRec, const int* __restrict__ ptOff, int* __restrict__ ptCursor,
                           unsigned long long* __restrict__ entKey, int* __restrict__ entA, int* __restrict__ entPt) {
  int A = min(counters[2], actCap);  // an overflowed pass is repeated by the host
  for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < A; a += gridDim.x * blockDim.x) {
    int t = actRec[a];
    int id[4];
    record_ids(recs[t], id);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (id[q] >= nSim) continue;
      int slot = ptOff[id[q]] + atomicAdd(ptCursor + id[q], 1);
      entKey[slot] = ((unsigned long long)t << 2) | (unsigned)q;
      entA[slot] = a;
      entPt[slot] = id[q];
    }
  }
}

// Each entry's rank among its point's entries (keys are distinct: a record
// names four distinct points) places it in key order: entSorted = (a << 2 | corner).
__global__ void k_col_rank(int nSim, int cap, const int* __restrict__ ptOff, const unsigned long long* __restrict__ entKey,
                           const int* __restrict__ entA, const int* __restrict__ entPt,
                           unsigned long long* __restrict__ entSorted) {
  int E = min(ptOff[nSim], cap);
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
    int i = entPt[e];
    unsigned long long k = entKey[e];
    int o = ptOff[i], end = ptOff[i + 1], rank = 0;
    for (int q = o; q < end; q++) rank += entKey[q] < k ? 1 : 0;
    entSorted[o + rank] = ((unsigned long long)entA[e] << 2) | (k & 3ull);
  }
}

__global__ void k_col_gather(int nSim, int ld, const int* __restrict__ ptOff,
                             const unsigned long long* __restrict__ entSorted, const double* __restrict__ actW,
                             const double* __restrict__ actT, double* __restrict__ diag, double* __restrict__ rhs) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nSim) return;
  int o = ptOff[i], e = ptOff[i + 1];
  double d = 0.0, r[3] = {0.0, 0.0, 0.0};
  for (int p = o; p < e; p++) {
    unsigned long long v = entSorted[p];
    int a = (int)(v >> 2), q = (int)(v & 3ull);
    double w = actW[a];
    d += w;
#pragma unroll
    for (int k = 0; k < 3; k++) r[k] += w * actT[(int64_t)a * 12 + 3 * q + k];
  }
  diag[i] = d;
#pragma unroll
  for (int k = 0; k < 3; k++) rhs[k * ld + i] = r[k];
}

// ---- adjoint of the narrow phase (active sets and candidate records frozen)

// For every active record at X: with colDbar / colRbar the adjoints of the
// collision diagonal and rhs (rows of simulated points), accumulate into ax the
// gradient of  sum_q colDbar_q w + colRbar_q . (w target_q)  with respect to the
// simulated points' positions (forward-mode over the 12 coordinates).
__global__ void __launch_bounds__(64) k_adj_collision(int nAct, int ld, int nSim, const int4* __restrict__ recs,
                                                      const int* __restrict__ act, const double* __restrict__ X,
                                                      const double* __restrict__ cm, double dHat, double kappa,
                                                      double dt, const double* __restrict__ colDbar,
                                                      const double* __restrict__ colRbar, double* __restrict__ ax) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nAct) return;
  int4 r = recs[act[a]];
  int type = r.x < 0 ? 1 : 0;
  int id[4];
  if (!record_ids(r, id)) return;
  P3 x[4];
#pragma unroll
  for (int q = 0; q < 4; q++) x[q] = load3(X, ld, id[q]);
  if (!cd_filter(type, x, dHat)) return;
  // activity in double first (the same test the step made)
  double cmq[4] = {cm[id[0]], cm[id[1]], cm[id[2]], cm[id[3]]};
  {
    double w;
    P3 tgt[4];
    if (!pair_terms(type, x, cmq, dHat, kappa, dt, &w, tgt)) return;
  }
  double gd[4], gr[4][3];
  bool any = false;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    bool simq = id[q] < nSim;
    gd[q] = simq ? colDbar[id[q]] : 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) gr[q][k] = simq ? colRbar[k * ld + id[q]] : 0.0;
    any = any || gd[q] != 0.0 || gr[q][0] != 0.0 || gr[q][1] != 0.0 || gr[q][2] != 0.0;
  }
  if (!any) return;
  using D12 = Dual<12>;
  P3T<D12> xd[4];
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int k = 0; k < 3; k++) xd[q].v[k] = D12::seed(x[q].v[k], 3 * q + k);
  D12 w;
  P3T<D12> tgt[4];
  pair_terms(type, xd, cmq, dHat, kappa, dt, &w, tgt);
  D12 L = D12(0.0);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    D12 acc = D12(gd[q]);
#pragma unroll
    for (int k = 0; k < 3; k++) acc += D12(gr[q][k]) * tgt[q].v[k];
    L += w * acc;  // sum_q gd_q w + gr_q . (w tgt_q)
  }
#pragma unroll
  for (int q = 0; q < 4; q++)
    if (id[q] < nSim)
#pragma unroll
      for (int k = 0; k < 3; k++) atomicAdd(ax + k * ld + id[q], L.d[3 * q + k]);
}

// Per-record TOI of the taped step (2.0 where a record is skipped or filtered),
// so that the record attaining the minimum can be found afterwards.
__global__ void __launch_bounds__(kBlock) k_ccd_records(int nRecs, int ld, const int4* __restrict__ recs,
                                                        const double* __restrict__ X0, const double* __restrict__ X1,
                                                        double filter, int iters, double step, double gapf,
                                                        int touchPolicy, double* __restrict__ recToi,
                                                        unsigned long long* __restrict__ toiBits) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int4 r = t < nRecs ? recs[t] : make_int4(0, 0, 0, 0);
  int type = r.x < 0 ? 1 : 0;
  int id[4];
  double toi = 2.0;
  if (t < nRecs && record_ids(r, id)) {
    P3 x[4], dx[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      x[q] = load3(X0, ld, id[q]);
      P3 y = load3(X1, ld, id[q]);
      dx[q] = sub(y, x[q]);
    }
    bool pass = type == 0 ? pt_ccd_filter(x, dx, filter) : ee_ccd_filter(x, dx, filter);
    if (pass) {
      double v = accd(type, x, dx, iters, step, gapf, touchPolicy);
      if (v < 2.0) toi = v;  // NaN stays 2
    }
  }
  if (t < nRecs) recToi[t] = toi;
  double mn = toi;
  for (int o = kWave / 2; o > 0; o >>= 1) {
    double other = __shfl_xor(mn, o, kWave);
    mn = other < mn ? other : mn;
  }
  if ((threadIdx.x % kWave) == 0 && mn < 1.0) atomicMin(toiBits, (unsigned long long)__double_as_longlong(mn));
}

// first record whose TOI equals the device minimum (only when the minimum < 1)
__global__ void k_toi_argmin(int nRecs, const double* __restrict__ recToi, const unsigned long long* __restrict__ toiBits,
                             int* __restrict__ argmin) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nRecs) return;
  double tmin = __longlong_as_double((long long)*toiBits);
  if (tmin < 1.0 && recToi[t] == tmin) atomicMin(argmin, t);
}

// ACCD exit path of the record argmin[0] (written to argmin[5]; -1 when there
// is none): one thread re-runs that record's ACCD
__global__ void k_argmin_path(int nRecs, int ld, const int4* __restrict__ recs, const double* __restrict__ X0,
                              const double* __restrict__ X1, double filter, int iters, double step, double gapf,
                              int touchPolicy, int* __restrict__ argmin) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int t = argmin[0];
  int path = -1;
  int id[4];
  if (t >= 0 && t < nRecs && record_ids(recs[t], id)) {
    const int type = recs[t].x < 0 ? 1 : 0;
    P3 x[4], dx[4];
    for (int q = 0; q < 4; q++) {
      x[q] = load3(X0, ld, id[q]);
      dx[q] = sub(load3(X1, ld, id[q]), x[q]);
    }
    if (type == 0 ? pt_ccd_filter(x, dx, filter) : ee_ccd_filter(x, dx, filter))
      accd(type, x, dx, iters, step, gapf, touchPolicy, &path);
  }
  argmin[5] = path;
}

// order-independent hash of the active barrier records (the sum of a 64-bit
// mix of each record index), added to *hash
__global__ void k_active_hash(const int* __restrict__ counters, int actCap, const int* __restrict__ actRec,
                              unsigned long long* __restrict__ hash) {
  const int A = min(counters[2], actCap);
  unsigned long long sum = 0;
  for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < A; a += gridDim.x * blockDim.x) {
    unsigned long long z = (unsigned long long)actRec[a] + 0x9e3779b97f4a7c15ull;  // splitmix64 finaliser
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    sum += z ^ (z >> 31);
  }
  for (int o = kWave / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o, kWave);
  if ((threadIdx.x % kWave) == 0 && sum) atomicAdd(hash, sum);
}

// d toi / d (x0, x1) of one record, scaled by g, added to ax0 / ax1 (simulated
// points): thread c of 24 differentiates along input coordinate c
__global__ void k_adj_ccd(int ld, int nSim, const int4* __restrict__ recs, int argmin, const double* __restrict__ X0,
                          const double* __restrict__ X1, double filter, int iters, double step, double gapf,
                          int touchPolicy, const double* __restrict__ gscale, double* __restrict__ ax0,
                          double* __restrict__ ax1) {
  int c = threadIdx.x;
  if (blockIdx.x != 0 || c >= 24 || argmin < 0) return;
  double g = *gscale;
  if (g == 0.0) return;
  int4 r = recs[argmin];
  int type = r.x < 0 ? 1 : 0;
  int id[4];
  record_ids(r, id);
  using D1 = Dual<1>;
  P3T<D1> x[4], dx[4];
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      // X1 holds the simulated rows only: static points end where they start
      const double* src1 = id[q] < nSim ? X1 : X0;
      x[q].v[k] = c == 3 * q + k ? D1::seed(X0[k * ld + id[q]], 0) : D1(X0[k * ld + id[q]]);
      D1 y = c == 12 + 3 * q + k ? D1::seed(src1[k * ld + id[q]], 0) : D1(src1[k * ld + id[q]]);
      dx[q].v[k] = y - x[q].v[k];
    }
  D1 toc = accd(type, x, dx, iters, step, gapf, touchPolicy);
  int q = (c % 12) / 3, k = c % 3;
  if (id[q] < nSim) {
    double* dst = c < 12 ? ax0 : ax1;
    dst[k * ld + id[q]] += g * toc.d[0];
  }
}

}  // namespace

void launch_toi_init(const DeviceMesh& m, hipStream_t s) {
  hipLaunchKernelGGL(k_toi_init, dim3(1), dim3(1), 0, s, m.toiBits);
}

void launch_ccd(const DeviceMesh& m, int nRecs, double filter, int iters, double step, double gap, int touchPolicy,
                hipStream_t s) {
  if (nRecs <= 0) return;
  launch_t(m.timer, TC_CCD, k_ccd, dim3((nRecs + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nRecs, m.ld, m.recs,
           m.Xold, m.X, filter, iters, step, gap, touchPolicy, m.toiBits);
}

void launch_ccd_blockmin(const DeviceMesh& m, int nRecs, const double* x0Aos, const double* x1Aos, double filter,
                         int iters, double step, double gap, int touchPolicy, double* blockMin, hipStream_t s) {
  if (nRecs <= 0) return;
  hipLaunchKernelGGL(k_ccd_blockmin, dim3((nRecs + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nRecs, m.recs, x0Aos,
                     x1Aos, filter, iters, step, gap, touchPolicy, blockMin);
}

void launch_ccd_records(const DeviceMesh& m, const int4* recs, int nRecs, const double* x0, const double* x1,
                        double filter, int iters, double step, double gap, int touchPolicy, double* recToi,
                        int* argmin, hipStream_t s) {
  if (nRecs <= 0) return;
  hipLaunchKernelGGL(k_ccd_records, dim3((nRecs + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nRecs, m.ld, recs, x0, x1,
                     filter, iters, step, gap, touchPolicy, recToi, m.toiBits);
  hipLaunchKernelGGL(k_toi_argmin, dim3((nRecs + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nRecs, recToi, m.toiBits,
                     argmin);
}

void launch_argmin_path(const DeviceMesh& m, int nRecs, const double* x0, const double* x1, double filter, int iters,
                        double step, double gap, int touchPolicy, int* argmin, hipStream_t s) {
  hipLaunchKernelGGL(k_argmin_path, dim3(1), dim3(64), 0, s, nRecs, m.ld, m.recs, x0, x1, filter, iters, step, gap,
                     touchPolicy, argmin);
}

void launch_active_hash(const DeviceMesh& m, unsigned long long* hash, hipStream_t s) {
  hipLaunchKernelGGL(k_active_hash, dim3(64), dim3(kBlock), 0, s, m.counters, m.actCap, m.actRec, hash);
}

void launch_adj_collision(const DeviceMesh& m, const int4* recs, const int* act, int nAct, const double* x, double dhat,
                          double kappa, double dt, const double* colDbar, const double* colRbar, double* ax,
                          hipStream_t s) {
  if (nAct <= 0) return;
  hipLaunchKernelGGL(k_adj_collision, dim3((nAct + 63) / 64), dim3(64), 0, s, nAct, m.ld, m.nSim, recs, act, x,
                     m.colMass, dhat, kappa, dt, colDbar, colRbar, ax);
}

void launch_adj_ccd(const DeviceMesh& m, const int4* recs, int argmin, const double* x0, const double* x1,
                    double filter, int iters, double step, double gap, int touchPolicy, const double* gscale,
                    double* ax0, double* ax1, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_ccd, dim3(1), dim3(64), 0, s, m.ld, m.nSim, recs, argmin, x0, x1, filter, iters, step, gap,
                     touchPolicy, gscale, ax0, ax1);
}

int launch_collision(DeviceMesh& m, int nRecs, double dhat, double kappa, double dt, hipStream_t s) {
  hipLaunchKernelGGL(k_col_clear, dim3((m.nSim + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, s, m.nSim, m.counters,
                     m.ptCount, m.ptCursor);
  if (nRecs > 0)
    launch_t(m.timer, TC_BARRIER, k_col_terms, dim3((nRecs + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nRecs, m.ld,
             m.nSim, m.recs, m.X, m.colMass, dhat, kappa, dt, m.actCap, m.actRec, m.actW, m.actT, m.ptCount,
             m.counters);
  launch_exclusive_scan(m.ptCount, m.ptOff, m.nSim + 1, nullptr, m.ptScan, s);
  launch_t(m.timer, TC_BARRIER, k_col_fill, dim3(512), dim3(kBlock), 0, s, m.ld, m.nSim, m.actCap, m.recs, m.counters,
           m.actRec, m.ptOff, m.ptCursor, m.entKey, m.entA, m.entPt);
  launch_t(m.timer, TC_BARRIER, k_col_rank, dim3(512), dim3(kBlock), 0, s, m.nSim, 4 * m.actCap, m.ptOff, m.entKey,
           m.entA, m.entPt, m.entSorted);
  launch_t(m.timer, TC_BARRIER, k_col_gather, dim3((m.nSim + kBlock - 1) / kBlock), dim3(kBlock), 0, s, m.nSim, m.ld,
           m.ptOff, m.entSorted, m.actW, m.actT, m.colDiag, m.colRhs);
  return 0;
}

}  // namespace pdipc
