// The PD local step (doLocalProjection, Simulator/PD-IPC/PdIpcSimulator.cpp:1624-1681)
// as two gathers instead of the reference's float atomics:
//   1. k_cloth_curvature: per cloth vertex, the Laplace-Beltrami vector a_i,
//      rescaled to the rest mean-curvature norm (PdIpcGpuFunc.cu:668-707);
//      k_tet_project: per tet, the corotational strain projection
//      w * U clamp(S) V^T Dm^{-T} G (PdIpcGpuFunc.cu:439-599) into 12 outputs.
//   2. k_rhs: per point, rhs = rhs0 + tet outputs + springs + bending gathered
//      from the 1-ring + collision rhs + position term.
// Every point's sum is formed by a few lanes (entries dealt round-robin)
// combined by a fixed shuffle tree, so the FP64 result is bit-reproducible.  T is the arithmetic type of the projections
// (double, or float in the FP32-split mode); edge vectors are formed in double
// and rounded once, so FP32 never sees absolute coordinates.
#include "device.hpp"

namespace pdipc {
namespace {

constexpr int kBlock = 256;
inline int grid_for(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

template <class T> __device__ inline T dsqrt(T x);
template <> __device__ inline double dsqrt<double>(double x) { return sqrt(x); }
template <> __device__ inline float dsqrt<float>(float x) { return sqrtf(x); }

// 3x3 SVD: cyclic Jacobi on F^T F with exact Givens rotations (run until the
// off-diagonal mass is negligible), eigenvalues sorted descending, V made a
// rotation, then Gram-Schmidt on F V gives U (a rotation) and the singular
// values, the last one carrying the sign of det F -- the output convention of
// McAdams et al. 2011 (the method of CudaSVD.cu:291-316).
template <class T>
__device__ inline void svd3(const T F[9], T U[9], T sig[3], T V[9]) {
  T a00 = F[0] * F[0] + F[3] * F[3] + F[6] * F[6];
  T a01 = F[0] * F[1] + F[3] * F[4] + F[6] * F[7];
  T a02 = F[0] * F[2] + F[3] * F[5] + F[6] * F[8];
  T a11 = F[1] * F[1] + F[4] * F[4] + F[7] * F[7];
  T a12 = F[1] * F[2] + F[4] * F[5] + F[7] * F[8];
  T a22 = F[2] * F[2] + F[5] * F[5] + F[8] * F[8];
  T W[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};  // row-major, columns are eigenvectors
  const int maxSweeps = sizeof(T) == 8 ? 12 : 6;
  const T tiny = sizeof(T) == 8 ? (T)1e-34 : (T)1e-14;
  for (int sw = 0; sw < maxSweeps; sw++) {
    T off = a01 * a01 + a12 * a12 + a02 * a02;
    T dia = a00 * a00 + a11 * a11 + a22 * a22;
    if (!(off > dia * tiny)) break;
    // rotate (p,q) = (0,1), (1,2), (0,2)
#pragma unroll
    for (int r = 0; r < 3; r++) {
      T *app, *aqq, *apq, *apr, *aqr;
      int p, q;
      if (r == 0) { app = &a00; aqq = &a11; apq = &a01; apr = &a02; aqr = &a12; p = 0; q = 1; }
      else if (r == 1) { app = &a11; aqq = &a22; apq = &a12; apr = &a01; aqr = &a02; p = 1; q = 2; }
      else { app = &a00; aqq = &a22; apq = &a02; apr = &a01; aqr = &a12; p = 0; q = 2; }
      T x = *apq;
      if (x != (T)0) {
        T tau = (*aqq - *app) / (2 * x);
        T t = (tau >= 0 ? (T)1 : (T)-1) / (fabs(tau) + dsqrt<T>((T)1 + tau * tau));
        T c = (T)1 / dsqrt<T>((T)1 + t * t), s = t * c;
        T pp = *app, qq = *aqq;
        *app = pp - t * x;
        *aqq = qq + t * x;
        *apq = 0;
        T pr = *apr, qr = *aqr;  // the entries coupling p and q to the third index
        *apr = c * pr - s * qr;
        *aqr = s * pr + c * qr;
#pragma unroll
        for (int k = 0; k < 3; k++) {
          T wp = W[3 * k + p], wq = W[3 * k + q];
          W[3 * k + p] = c * wp - s * wq;
          W[3 * k + q] = s * wp + c * wq;
        }
      }
    }
  }
  T lam[3] = {a00, a11, a22};
  int o0 = 0, o1 = 1, o2 = 2, tmp;
  if (lam[o1] > lam[o0]) { tmp = o0; o0 = o1; o1 = tmp; }
  if (lam[o2] > lam[o0]) { tmp = o0; o0 = o2; o2 = tmp; }
  if (lam[o2] > lam[o1]) { tmp = o1; o1 = o2; o2 = tmp; }
  // columns reordered with selects (a run-time index into W would put it in
  // scratch memory)
  const int ord[3] = {o0, o1, o2};
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int c = 0; c < 3; c++) V[3 * k + c] = ord[c] == 0 ? W[3 * k] : (ord[c] == 1 ? W[3 * k + 1] : W[3 * k + 2]);
  T detV = V[0] * (V[4] * V[8] - V[5] * V[7]) - V[1] * (V[3] * V[8] - V[5] * V[6]) + V[2] * (V[3] * V[7] - V[4] * V[6]);
  if (detV < 0) { V[2] = -V[2]; V[5] = -V[5]; V[8] = -V[8]; }
  T b[3][3];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int r = 0; r < 3; r++) b[c][r] = F[3 * r] * V[c] + F[3 * r + 1] * V[3 + c] + F[3 * r + 2] * V[6 + c];
  T u0[3], u1[3], u2[3];
  T s0 = dsqrt<T>(b[0][0] * b[0][0] + b[0][1] * b[0][1] + b[0][2] * b[0][2]);
  if (s0 > (T)1e-30) { u0[0] = b[0][0] / s0; u0[1] = b[0][1] / s0; u0[2] = b[0][2] / s0; }
  else { u0[0] = 1; u0[1] = 0; u0[2] = 0; }
  T pr = u0[0] * b[1][0] + u0[1] * b[1][1] + u0[2] * b[1][2];
  u1[0] = b[1][0] - pr * u0[0]; u1[1] = b[1][1] - pr * u0[1]; u1[2] = b[1][2] - pr * u0[2];
  T s1 = dsqrt<T>(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
  if (s1 > (T)1e-30 * (s0 > 1 ? s0 : (T)1)) {
    u1[0] /= s1; u1[1] /= s1; u1[2] /= s1;
  } else {
    T ax = fabs(u0[0]), ay = fabs(u0[1]), az = fabs(u0[2]);
    const int m = (ax < ay) ? (ax < az ? 0 : 2) : (ay < az ? 1 : 2);
    const T t[3] = {m == 0 ? (T)1 : (T)0, m == 1 ? (T)1 : (T)0, m == 2 ? (T)1 : (T)0};
    u1[0] = u0[1] * t[2] - u0[2] * t[1];
    u1[1] = u0[2] * t[0] - u0[0] * t[2];
    u1[2] = u0[0] * t[1] - u0[1] * t[0];
    T n = dsqrt<T>(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
    u1[0] /= n; u1[1] /= n; u1[2] /= n;
    s1 = 0;
  }
  u2[0] = u0[1] * u1[2] - u0[2] * u1[1];
  u2[1] = u0[2] * u1[0] - u0[0] * u1[2];
  u2[2] = u0[0] * u1[1] - u0[1] * u1[0];
  T s2 = u2[0] * b[2][0] + u2[1] * b[2][1] + u2[2] * b[2][2];
#pragma unroll
  for (int r = 0; r < 3; r++) { U[3 * r] = u0[r]; U[3 * r + 1] = u1[r]; U[3 * r + 2] = u2[r]; }
  sig[0] = s0; sig[1] = s1; sig[2] = s2;
}

template <class T>
__global__ void k_cloth_curvature(int nCloth, int ld, const int* __restrict__ cCenter, const int* __restrict__ cOff,
                                  const int* __restrict__ cNbr, const double* __restrict__ cLbo,
                                  const double* __restrict__ cH0, const double* __restrict__ X,
                                  double* __restrict__ A) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nCloth) return;
  int i = cCenter[c];
  int o = cOff[c], e = cOff[c + 1];
  double xi[3] = {X[i], X[ld + i], X[2 * ld + i]};
  // a = sum_s L_s x_s = sum_{s>0} L_s (x_s - x_i) + (sum_s L_s) x_i  (the row sum is ~0)
  double lsum = cLbo[o];
  T a[3] = {0, 0, 0};
  for (int s = o + 1; s < e; s++) {
    int j = cNbr[s];
    T l = (T)cLbo[s];
    lsum += cLbo[s];
#pragma unroll
    for (int k = 0; k < 3; k++) a[k] += l * (T)(X[k * ld + j] - xi[k]);
  }
  double ad[3];
#pragma unroll
  for (int k = 0; k < 3; k++) ad[k] = (double)a[k] + lsum * xi[k];
  double n = sqrt(ad[0] * ad[0] + ad[1] * ad[1] + ad[2] * ad[2]);
  if (n >= 1e-10) {
    double sc = cH0[c] / n;
#pragma unroll
    for (int k = 0; k < 3; k++) ad[k] *= sc;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) A[k * ld + i] = ad[k];
}

template <class T>
__global__ void k_tet_project(int nT, int ld, T sigLo, T sigHi, const int4* __restrict__ tets,
                              const double* __restrict__ dminv, const double* __restrict__ w,
                              const double* __restrict__ X, double* __restrict__ out) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nT) return;
  int4 q = tets[t];
  int id[4] = {q.x, q.y, q.z, q.w};
  T Ds[9];
#pragma unroll
  for (int r = 0; r < 3; r++) {
    double x0 = X[r * ld + id[0]];
#pragma unroll
    for (int c = 0; c < 3; c++) Ds[3 * r + c] = (T)(X[r * ld + id[c + 1]] - x0);
  }
  T idm[9];
#pragma unroll
  for (int k = 0; k < 9; k++) idm[k] = (T)dminv[(int64_t)k * nT + t];
  T F[9];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) F[3 * r + c] = Ds[3 * r] * idm[c] + Ds[3 * r + 1] * idm[3 + c] + Ds[3 * r + 2] * idm[6 + c];
  T U[9], s[3], V[9];
  svd3<T>(F, U, s, V);
  if (s[0] * s[1] * s[2] < 0) {  // inverted: flip the last value and re-sort the values only
    s[2] = -s[2];
    T hi = s[0], mid = s[1], lo = s[2], tt;
    if (mid < lo) { tt = lo; lo = mid; mid = tt; }
    if (hi < lo) { tt = lo; lo = hi; hi = tt; }
    if (hi < mid) { tt = mid; mid = hi; hi = tt; }
    s[0] = hi; s[1] = mid; s[2] = lo;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    T r = s[k] < sigHi ? s[k] : sigHi;
    s[k] = r < sigLo ? sigLo : r;
  }
  T P[9];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) P[3 * r + c] = U[3 * r] * s[0] * V[3 * c] + U[3 * r + 1] * s[1] * V[3 * c + 1] + U[3 * r + 2] * s[2] * V[3 * c + 2];
  // result = w * P * [ -Dm^{-T} 1 | Dm^{-T} ]   (PdIpcGpuFunc.cu:554-580)
  T wt = (T)w[t];
#pragma unroll
  for (int r = 0; r < 3; r++) {
    T g1 = P[3 * r] * idm[0] + P[3 * r + 1] * idm[1] + P[3 * r + 2] * idm[2];
    T g2 = P[3 * r] * idm[3] + P[3 * r + 1] * idm[4] + P[3 * r + 2] * idm[5];
    T g3 = P[3 * r] * idm[6] + P[3 * r + 1] * idm[7] + P[3 * r + 2] * idm[8];
    T g0 = -(g1 + g2 + g3);
    out[(int64_t)(0 * 3 + r) * nT + t] = (double)(g0 * wt);
    out[(int64_t)(1 * 3 + r) * nT + t] = (double)(g1 * wt);
    out[(int64_t)(2 * 3 + r) * nT + t] = (double)(g2 * wt);
    out[(int64_t)(3 * 3 + r) * nT + t] = (double)(g3 * wt);
  }
}

// RES (FP32-split): also the first pass of the solve's FP64 residual for the
// current X and D (as k_res_a in k_solve.hip: v = M x + w D^{-1} b and the
// pointwise part (1 - w) M x + (2 - w) w D^{-1} b - x), saving a launch.
struct ResidualArgs {
  double w;
  const int* off;
  const int* col;
  const double* val;
  const double* D;
  Row64* v;
  Row64* pt;
};

// kRhsLanes lanes per point: ring slots, tet corners and (RES) R1 entries are
// dealt round-robin and the partial sums combined by a fixed shuffle tree.
// A lane fetches kRhsBatch of its tet corners (kResBatch R1 entries) at a time, all
// loads first, then adds them in order, so that several memory round trips
// are in flight per lane; the sums are those of the one-at-a-time loop.
constexpr int kRhsLanes = 4;
constexpr int kRhsBatch = 4;
// R1 entries per lane fetched together in the residual pass (RES): a cloth row
// has about 19, so one batch of 6 over the 4 lanes covers it (14.5 -> 13.7 us
// for k_rhs<float, true> at 65,536 points against batches of 4)
constexpr int kResBatch = 6;

template <class V>
__device__ inline V rhs_lanes_sum(V v) {
#pragma unroll
  for (int o = kRhsLanes / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// ring slots of each lane fetched together (all loads first, then the sums
// in slot order, so the sums are those of the one-at-a-time loop)
constexpr int kRingBatch = 2;

template <class T, bool RES>
__global__ void k_rhs(int nSim, int ld, int nT, int withCol, ResidualArgs ra, const double* __restrict__ initRhs,
                      const int* __restrict__ vtOff, const int* __restrict__ vtList, const double* __restrict__ tetOut,
                      const int2* __restrict__ ringRange, const int* __restrict__ cNbr,
                      const double* __restrict__ cRest, const double* __restrict__ cStrain,
                      const double* __restrict__ cBendIn, const double* __restrict__ A, const double* __restrict__ X,
                      const double* __restrict__ colRhs, const double* __restrict__ fixK,
                      const double* __restrict__ Xrest, double* __restrict__ rhs) {
  int gt = blockIdx.x * blockDim.x + threadIdx.x;
  int i = gt / kRhsLanes, sub = gt % kRhsLanes;
  bool live = i < nSim;
  const int ii = live ? i : nSim - 1;
  // every operand that depends on the point alone, up front
  const int2 ring = ringRange[ii];  // the point's cloth ring slots (empty for tet points)
  const int q0 = live ? vtOff[ii] : 0, q1 = live ? vtOff[ii + 1] : 0;
  const int r0 = RES ? ra.off[ii] : 0, r1 = RES && live ? ra.off[ii + 1] : 0;
  const double xi[3] = {X[ii], X[ld + ii], X[2 * ld + ii]};
  double init[3] = {0, 0, 0}, col[3] = {0, 0, 0}, kf = 0.0, di = 1.0;
  if (sub == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      init[k] = initRhs[k * ld + ii];
      if (withCol) col[k] = colRhs[k * ld + ii];
    }
    kf = fixK[ii];
    if (RES) di = ra.D[ii];
  }
  double tr[3] = {0, 0, 0};  // tet outputs
  for (int q = q0 + sub; q < q1; q += kRhsBatch * kRhsLanes) {
    int tc[kRhsBatch];
#pragma unroll
    for (int b = 0; b < kRhsBatch; b++) tc[b] = q + b * kRhsLanes < q1 ? vtList[q + b * kRhsLanes] : -1;
    double a[kRhsBatch][3];
#pragma unroll
    for (int b = 0; b < kRhsBatch; b++)
#pragma unroll
      for (int k = 0; k < 3; k++)
        a[b][k] = tc[b] >= 0 ? tetOut[(int64_t)((tc[b] & 3) * 3 + k) * nT + (tc[b] >> 2)] : 0.0;
#pragma unroll
    for (int b = 0; b < kRhsBatch; b++)
      if (tc[b] >= 0)
#pragma unroll
        for (int k = 0; k < 3; k++) tr[k] += a[b][k];
  }
  // cloth ring: springs over the neighbour slots, the bending gather over all
  // slots (the first slot is the point itself), in one pass
  T sp[3] = {0, 0, 0};
  double bd[3] = {0, 0, 0};
  const int o = live ? ring.x : 0, e = live ? ring.y : 0;
  for (int s0 = o + sub; s0 < e; s0 += kRingBatch * kRhsLanes) {
    int j[kRingBatch];
    double xj[kRingBatch][3], aj[kRingBatch][3], bw[kRingBatch], rest[kRingBatch], stiff[kRingBatch];
#pragma unroll
    for (int b = 0; b < kRingBatch; b++) {
      const int sl = s0 + b * kRhsLanes;
      const bool in = sl < e;
      j[b] = in ? cNbr[sl] : ii;
      bw[b] = in ? cBendIn[sl] : 0.0;
      rest[b] = in && sl > o ? cRest[sl] : 0.0;
      stiff[b] = in && sl > o ? cStrain[sl] : 0.0;
    }
#pragma unroll
    for (int b = 0; b < kRingBatch; b++)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        xj[b][k] = X[k * ld + j[b]];
        aj[b][k] = A[k * ld + j[b]];
      }
#pragma unroll
    for (int b = 0; b < kRingBatch; b++) {
      const int sl = s0 + b * kRhsLanes;
      if (sl >= e) continue;
      if (sl > o) {
        T d[3];
#pragma unroll
        for (int k = 0; k < 3; k++) d[k] = (T)(xi[k] - xj[b][k]);
        T len = dsqrt<T>(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        T rc = (T)stiff[b] * ((T)rest[b] / len);
#pragma unroll
        for (int k = 0; k < 3; k++) sp[k] += rc * d[k];
      }
#pragma unroll
      for (int k = 0; k < 3; k++) bd[k] += bw[b] * aj[b][k];
    }
  }
  double sx[3] = {0, 0, 0};
  if (RES) {
    for (int q = r0 + sub; q < r1; q += kResBatch * kRhsLanes) {
      int j[kResBatch];
      double a[kResBatch], xv[kResBatch][3];
#pragma unroll
      for (int b = 0; b < kResBatch; b++) {
        const bool in = q + b * kRhsLanes < r1;
        j[b] = in ? ra.col[q + b * kRhsLanes] : ii;
        a[b] = in ? ra.val[q + b * kRhsLanes] : 0.0;
      }
#pragma unroll
      for (int b = 0; b < kResBatch; b++)
#pragma unroll
        for (int k = 0; k < 3; k++) xv[b][k] = X[k * ld + j[b]];
#pragma unroll
      for (int b = 0; b < kResBatch; b++)
        if (q + b * kRhsLanes < r1)
#pragma unroll
          for (int k = 0; k < 3; k++) sx[k] += a[b] * xv[b][k];
    }
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    tr[k] = rhs_lanes_sum(tr[k]);
    sp[k] = rhs_lanes_sum(sp[k]);
    bd[k] = rhs_lanes_sum(bd[k]);
    if (RES) sx[k] = rhs_lanes_sum(sx[k]);
  }
  if (!live || sub != 0) return;
  double r[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    r[k] = init[k] + tr[k];  // the reference's order: rhs0, tets, cloth, collision, positions
    if (e > o) r[k] += (double)sp[k] + bd[k];
    if (withCol) r[k] += col[k];
  }
  if (kf != 0.0) {
#pragma unroll
    for (int k = 0; k < 3; k++) r[k] += kf * Xrest[k * ld + i];
  }
#pragma unroll
  for (int k = 0; k < 3; k++) rhs[k * ld + i] = r[k];
  if (RES) {
    const double w = ra.w;
    double vv[3], pp[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double y = (1.0 - w) * xi[k] + w * sx[k] / di;
      double u = r[k] / di;
      vv[k] = y + w * u;
      pp[k] = (1.0 - w) * y + (2.0 - w) * w * u - xi[k];
    }
    ra.v[i] = Row64{vv[0], vv[1], vv[2], 0.0};
    ra.pt[i] = Row64{pp[0], pp[1], pp[2], 0.0};
  }
}

template <class T>
void local_impl(const DeviceMesh& m, T lo, T hi, int withCol, double resW, hipStream_t s) {
  if (m.nCloth > 0)
    launch_t(m.timer, TC_PROJECT, k_cloth_curvature<T>, dim3(grid_for(m.nCloth)), dim3(kBlock), 0, s, m.nCloth, m.ld,
             m.cCenter, m.cOff, m.cNbr, m.cLbo, m.cH0, m.X, m.clothA);
  if (m.nTets > 0)
    launch_t(m.timer, TC_PROJECT, k_tet_project<T>, dim3(grid_for(m.nTets)), dim3(kBlock), 0, s, m.nTets, m.ld, lo, hi,
             m.tetIdx, m.tetDminv, m.tetW, m.X, m.tetOut);
  const ResidualArgs ra{resW, m.r1Off, m.r1Col, m.r1Val, m.sysDiag, m.xa0, m.cA};
  const dim3 grid(grid_for((int64_t)m.nSim * kRhsLanes));
  launch_t(m.timer, TC_RHS, resW >= 0.0 ? k_rhs<T, true> : k_rhs<T, false>, grid, dim3(kBlock), 0, s, m.nSim, m.ld,
           m.nTets, withCol, ra, m.initRhs, m.vtOff, m.vtList, m.tetOut, m.ringRange, m.cNbr, m.cRest, m.cStrain,
           m.cBendIn, m.clothA, m.X, m.colRhs, m.fixK, m.Xrest, m.sysRhs);
}

// ---- adjoint of the local step (FP64), evaluated at the recorded x

// cloth: abar_i = sum_slots bend(i -> slot) bbar_slot, pulled back through the
// rescale a -> a H0 / |a| (Jacobian (H0/|a|)(I - a_hat a_hat^T), symmetric)
__global__ void k_adj_cloth_a(int nCloth, int ld, const int* __restrict__ cCenter, const int* __restrict__ cOff,
                              const int* __restrict__ cNbr, const double* __restrict__ cLbo,
                              const double* __restrict__ cH0, const double* __restrict__ cBendOut,
                              const double* __restrict__ X, const double* __restrict__ bbar,
                              double* __restrict__ Aadj) {
  // kRhsLanes lanes per cloth vertex share its ring slots (lanes past nCloth
  // run on the last vertex without writing)
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int sub = gt % kRhsLanes;
  const bool live = gt / kRhsLanes < nCloth;
  const int c = live ? gt / kRhsLanes : nCloth - 1;
  int i = cCenter[c];
  int o = cOff[c], e = cOff[c + 1];
  double ab[3] = {0, 0, 0};
  for (int s = o + sub; s < e; s += kRhsLanes) {
    int j = cNbr[s];
    double b = cBendOut[s];
#pragma unroll
    for (int k = 0; k < 3; k++) ab[k] += b * bbar[k * ld + j];
  }
  double xi[3] = {X[i], X[ld + i], X[2 * ld + i]};
  double lsum = sub == 0 ? cLbo[o] : 0.0;
  double a[3] = {0, 0, 0};
  for (int s = o + 1 + sub; s < e; s += kRhsLanes) {
    int j = cNbr[s];
    double l = cLbo[s];
    lsum += l;
#pragma unroll
    for (int k = 0; k < 3; k++) a[k] += l * (X[k * ld + j] - xi[k]);
  }
  lsum = rhs_lanes_sum(lsum);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    ab[k] = rhs_lanes_sum(ab[k]);
    a[k] = rhs_lanes_sum(a[k]);
  }
  if (!live || sub != 0) return;
#pragma unroll
  for (int k = 0; k < 3; k++) a[k] += lsum * xi[k];
  double n = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  if (n >= 1e-10) {
    double u[3] = {a[0] / n, a[1] / n, a[2] / n};
    double pr = u[0] * ab[0] + u[1] * ab[1] + u[2] * ab[2];
    double sc = cH0[c] / n;
#pragma unroll
    for (int k = 0; k < 3; k++) ab[k] = sc * (ab[k] - u[k] * pr);
  }
#pragma unroll
  for (int k = 0; k < 3; k++) Aadj[k * ld + i] = ab[k];
}

// Tets: Pbar = w sum_c bbar_c h_c^T, Fbar = dP(F)^T[Pbar], xbar_c = Fbar h_c.
// P(F) = U g(S) V^T with g the flip / re-sort / clamp of the singular values;
// with Ft = U^T dF V:  dPt_kk = sum_m dg_k/ds_m Ft_mm, and for i != j
//   dPt_ij = A (Ft_ij + Ft_ji)/2 + B (Ft_ij - Ft_ji)/2,
//   A = (g_j - g_i)/(s_j - s_i),  B = (g_i + g_j)/(s_i + s_j)
// (the differential of U diag(g) V^T through dU = U Omega_U, dV = V Omega_V);
// the adjoint swaps Pbar_ij, Pbar_ji in the same combination.
__global__ void __launch_bounds__(kBlock) k_adj_tet(int nT, int ld, double sigLo, double sigHi,
                                                   const int4* __restrict__ tets, const double* __restrict__ dminv,
                                                   const double* __restrict__ w, const double* __restrict__ X,
                                                   const double* __restrict__ bbar, double* __restrict__ out) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nT) return;
  int4 q = tets[t];
  int id[4] = {q.x, q.y, q.z, q.w};
  double Ds[9];
#pragma unroll
  for (int r = 0; r < 3; r++) {
    double x0 = X[r * ld + id[0]];
#pragma unroll
    for (int c = 0; c < 3; c++) Ds[3 * r + c] = X[r * ld + id[c + 1]] - x0;
  }
  double idm[9];
#pragma unroll
  for (int k = 0; k < 9; k++) idm[k] = dminv[(int64_t)k * nT + t];
  double F[9];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) F[3 * r + c] = Ds[3 * r] * idm[c] + Ds[3 * r + 1] * idm[3 + c] + Ds[3 * r + 2] * idm[6 + c];
  double U[9], s[3], V[9];
  svd3<double>(F, U, s, V);
  // g_k and dg_k/ds_{src[k]} as in the forward projection
  double sv[3] = {s[0], s[1], s[2]}, sgn[3] = {1.0, 1.0, 1.0};
  int src[3] = {0, 1, 2};
  if (s[0] * s[1] * s[2] < 0) {
    sv[2] = -sv[2];
    sgn[2] = -1.0;
    // the forward's three compare-swaps on (hi, mid, lo), tracking where each value came from
    int ih = 0, im = 1, il = 2;
    double hi = sv[0], mid = sv[1], lo = sv[2], tt;
    int ti;
    if (mid < lo) { tt = lo; lo = mid; mid = tt; ti = il; il = im; im = ti; }
    if (hi < lo) { tt = lo; lo = hi; hi = tt; ti = il; il = ih; ih = ti; }
    if (hi < mid) { tt = mid; mid = hi; hi = tt; ti = im; im = ih; ih = ti; }
    src[0] = ih; src[1] = im; src[2] = il;
  }
  double g[3], dg[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {  // selects, not run-time indices (which would use scratch memory)
    const int sk = src[k];
    double v = sk == 0 ? sv[0] : (sk == 1 ? sv[1] : sv[2]);
    double sg = sk == 0 ? sgn[0] : (sk == 1 ? sgn[1] : sgn[2]);
    double r = v < sigHi ? v : sigHi;
    g[k] = r < sigLo ? sigLo : r;
    dg[k] = (v < sigHi && !(v < sigLo)) ? sg : 0.0;  // d g_k / d s_{src[k]}
  }
  // Pbar
  double h[4][3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    h[1][k] = idm[k];
    h[2][k] = idm[3 + k];
    h[3][k] = idm[6 + k];
    h[0][k] = -(h[1][k] + h[2][k] + h[3][k]);
  }
  double wt = w[t];
  double bb[4][3];
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int r = 0; r < 3; r++) bb[c][r] = bbar[r * ld + id[c]];
  double Pb[9];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int k = 0; k < 3; k++)
      Pb[3 * r + k] = wt * (bb[0][r] * h[0][k] + bb[1][r] * h[1][k] + bb[2][r] * h[2][k] + bb[3][r] * h[3][k]);
  // Pt = U^T Pbar V
  double T1[9], Pt[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) T1[3 * i + k] = U[i] * Pb[k] + U[3 + i] * Pb[3 + k] + U[6 + i] * Pb[6 + k];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) Pt[3 * i + j] = T1[3 * i] * V[j] + T1[3 * i + 1] * V[3 + j] + T1[3 * i + 2] * V[6 + j];
  double Ft[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      if (src[k] == j) Ft[4 * j] += dg[k] * Pt[4 * k];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if (i == j) continue;
      double ds = s[j] - s[i], ss = s[i] + s[j];
      double scale = fabs(s[i]) + fabs(s[j]) + 1.0;
      double A = fabs(ds) > 1e-12 * scale ? (g[j] - g[i]) / ds : 0.5 * (fabs(dg[i]) + fabs(dg[j]));
      double B = fabs(ss) > 1e-12 * scale ? (g[i] + g[j]) / ss : 0.0;
      Ft[3 * i + j] = 0.5 * (A + B) * Pt[3 * i + j] + 0.5 * (A - B) * Pt[3 * j + i];
    }
  // Fbar = U Ft V^T
  double T2[9], Fb[9];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int j = 0; j < 3; j++) T2[3 * r + j] = U[3 * r] * Ft[j] + U[3 * r + 1] * Ft[3 + j] + U[3 * r + 2] * Ft[6 + j];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int k = 0; k < 3; k++) Fb[3 * r + k] = T2[3 * r] * V[3 * k] + T2[3 * r + 1] * V[3 * k + 1] + T2[3 * r + 2] * V[3 * k + 2];
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int r = 0; r < 3; r++)
      out[(int64_t)(c * 3 + r) * nT + t] = Fb[3 * r] * h[c][0] + Fb[3 * r + 1] * h[c][1] + Fb[3 * r + 2] * h[c][2];
}

// per point: ax = D u0 + (local step)^T bbar; spring-weight adjoints per slot;
// bbar also flows to rhs0 and (after the first outer iteration) the collision rhs
__global__ void k_adj_gather(int nSim, int ld, int nT, const double* __restrict__ D, const double* __restrict__ u0,
                             const int* __restrict__ vtOff, const int* __restrict__ vtList,
                             const double* __restrict__ tetAdj, const int* __restrict__ clothOfPoint,
                             const int* __restrict__ cOff, const int* __restrict__ cNbr,
                             const double* __restrict__ cRest, const double* __restrict__ cStrain,
                             const double* __restrict__ cLboIn, const double* __restrict__ Aadj,
                             const double* __restrict__ X, const double* __restrict__ bbar, double* __restrict__ ax,
                             double* __restrict__ sbar, double* __restrict__ rhs0bar, double* __restrict__ colRbar) {
  // kRhsLanes lanes per point share its tet corners and ring slots; lanes past
  // nSim run on the last point without writing, so every lane takes part in the
  // reduction; the point's own term enters once, on lane 0
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int sub = gt % kRhsLanes;
  const bool live = gt / kRhsLanes < nSim;
  const int i = live ? gt / kRhsLanes : nSim - 1;
  double di = D[i];
  double bi[3] = {bbar[i], bbar[ld + i], bbar[2 * ld + i]};
  double r[3] = {0.0, 0.0, 0.0};
  if (sub == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) r[k] = di * u0[k * ld + i];
  }
  for (int q = vtOff[i] + sub; q < vtOff[i + 1]; q += kRhsLanes) {
    int tc = vtList[q];
    int t = tc >> 2, corner = tc & 3;
#pragma unroll
    for (int k = 0; k < 3; k++) r[k] += tetAdj[(int64_t)(corner * 3 + k) * nT + t];
  }
  int c = clothOfPoint[i];
  if (c >= 0) {
    int o = cOff[c], e = cOff[c + 1];
    double xi[3] = {X[i], X[ld + i], X[2 * ld + i]};
    for (int s = o + 1 + sub; s < e; s += kRhsLanes) {
      int j = cNbr[s];
      double d[3], db[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        d[k] = xi[k] - X[k * ld + j];
        db[k] = bi[k] - bbar[k * ld + j];
      }
      double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      double n[3] = {d[0] / len, d[1] / len, d[2] / len};
      double coef = cStrain[s] * cRest[s] / len;
      double pr = n[0] * db[0] + n[1] * db[1] + n[2] * db[2];
#pragma unroll
      for (int k = 0; k < 3; k++) r[k] += coef * (db[k] - n[k] * pr);
      if (live) sbar[s] += cRest[s] * (n[0] * bi[0] + n[1] * bi[1] + n[2] * bi[2]);
    }
    for (int s = o + sub; s < e; s += kRhsLanes) {
      int j = cNbr[s];
      double l = cLboIn[s];
#pragma unroll
      for (int k = 0; k < 3; k++) r[k] += l * Aadj[k * ld + j];
    }
  }
#pragma unroll
  for (int k = 0; k < 3; k++) r[k] = rhs_lanes_sum(r[k]);
  if (!live || sub != 0) return;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    ax[k * ld + i] = r[k];
    rhs0bar[k * ld + i] += bi[k];
    if (colRbar) colRbar[k * ld + i] += bi[k];
  }
}

}  // namespace

void launch_adj_local(const DeviceMesh& m, const AdjMesh& a, double sigLo, double sigHi, const double* x,
                      const double* bbar, const double* u0, double* aA, double* tetAdj, double* ax, double* sbar,
                      double* rhs0bar, double* colRbar, hipStream_t s) {
  if (m.nCloth > 0)
    hipLaunchKernelGGL(k_adj_cloth_a, dim3(grid_for((int64_t)m.nCloth * kRhsLanes)), dim3(kBlock), 0, s, m.nCloth,
                       m.ld, m.cCenter, m.cOff, m.cNbr, m.cLbo, m.cH0, a.cBendOut, x, bbar, aA);
  if (m.nTets > 0)
    hipLaunchKernelGGL(k_adj_tet, dim3(grid_for(m.nTets)), dim3(kBlock), 0, s, m.nTets, m.ld, sigLo, sigHi, m.tetIdx,
                       m.tetDminv, m.tetW, x, bbar, tetAdj);
  hipLaunchKernelGGL(k_adj_gather, dim3(grid_for((int64_t)m.nSim * kRhsLanes)), dim3(kBlock), 0, s, m.nSim, m.ld,
                     m.nTets, m.sysDiag, u0, m.vtOff, m.vtList, tetAdj, m.clothOfPoint, m.cOff, m.cNbr, m.cRest,
                     m.cStrain, a.cLboIn, aA, x, bbar, ax, sbar, rhs0bar, colRbar);
}

void launch_local(const DeviceMesh& m, int precision, double sigLo, double sigHi, int withCollision, hipStream_t s,
                  double residualOmega) {
  if (precision == 1) local_impl<float>(m, (float)sigLo, (float)sigHi, withCollision, residualOmega, s);
  else local_impl<double>(m, sigLo, sigHi, withCollision, residualOmega, s);
}

}  // namespace pdipc
