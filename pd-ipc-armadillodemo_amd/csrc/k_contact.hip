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

namespace pdipc {
namespace {

constexpr int kBlock = 256;

struct P3 { double v[3]; };

__device__ inline P3 sub(const P3& a, const P3& b) { return {{a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]}}; }
__device__ inline double dot(const P3& a, const P3& b) { return a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2]; }
__device__ inline P3 cross(const P3& a, const P3& b) {
  return {{a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2], a.v[0] * b.v[1] - a.v[1] * b.v[0]}};
}

// VertexEdgeSqDistance (PdIpcGpuFunc.cu:2155-2174)
__device__ inline double ve_d2(const P3& v, const P3& ea, const P3& eb, double* r) {
  P3 va = sub(v, ea), ba = sub(eb, ea);
  double vb = dot(va, ba), bb = dot(ba, ba);
  *r = vb < 0 ? 0.0 : (vb > bb ? 1.0 : vb / bb);
  P3 d;
#pragma unroll
  for (int k = 0; k < 3; k++) d.v[k] = v.v[k] - (ea.v[k] * (1.0 - *r) + eb.v[k] * *r);
  return dot(d, d);
}

// EdgeEdgeSqDistance (PdIpcGpuFunc.cu:2176-2236); parallel edges (cross product
// below 1e-24) take the minimum over all four endpoint-edge distances.
__device__ inline double ee_d2(const P3& xi, const P3& xj, const P3& xa, const P3& xb, double* r, double* s, P3* dir) {
  P3 xba = sub(xb, xa), xji = sub(xj, xi), xai = sub(xa, xi);
  P3 n = cross(xji, xba);
  double nn = dot(n, n);
  double waiji = dot(n, cross(xai, xji));
  double waiba = dot(n, cross(xai, xba));
  if (nn > 1e-24 && waiji >= 0 && waiji <= nn && waiba >= 0 && waiba <= nn) {
    *r = waiba / nn;
    *s = waiji / nn;
  } else {
    bool dg = !(nn > 1e-24);
    double best = 999999999.0, d, v;
    *r = 0; *s = 0;
    if ((dg || waiba < 0) && (d = ve_d2(xi, xa, xb, &v)) < best) { best = d; *r = 0; *s = v; }
    if ((dg || waiba > nn) && (d = ve_d2(xj, xa, xb, &v)) < best) { best = d; *r = 1; *s = v; }
    if ((dg || waiji < 0) && (d = ve_d2(xa, xi, xj, &v)) < best) { best = d; *r = v; *s = 0; }
    if ((dg || waiji > nn) && (d = ve_d2(xb, xi, xj, &v)) < best) { best = d; *r = v; *s = 1; }
  }
#pragma unroll
  for (int k = 0; k < 3; k++) dir->v[k] = xi.v[k] * (1 - *r) + xj.v[k] * *r - xa.v[k] * (1.0 - *s) - xb.v[k] * *s;
  return dot(*dir, *dir);
}

// VertexTriangleDistance (PdIpcGpuFunc.cu:2238-2292); degenerate triangles take
// the minimum over the three edges.
__device__ inline double pt_d2(const P3& xi, const P3& xa, const P3& xb, const P3& xc, double* ba, double* bb,
                               double* bc, P3* dir) {
  P3 xba = sub(xb, xa), xca = sub(xc, xa), xia = sub(xi, xa);
  P3 n = cross(xba, xca);
  double nn = dot(n, n);
  double wiaca = dot(n, cross(xia, xca));
  double wbaia = dot(n, cross(xba, xia));
  if (nn > 1e-24 && wiaca >= 0 && wbaia >= 0 && nn - wiaca - wbaia >= 0) {
    *bb = wiaca / nn;
    *bc = wbaia / nn;
    *ba = 1 - *bb - *bc;
  } else {
    bool dg = !(nn > 1e-24);
    double best = 999999999.0, r, d;
    *ba = 1; *bb = 0; *bc = 0;
    if ((dg || nn - wiaca - wbaia < 0) && (d = ve_d2(xi, xb, xc, &r)) < best) { best = d; *bb = 1 - r; *bc = r; *ba = 0; }
    if ((dg || wiaca < 0) && (d = ve_d2(xi, xa, xc, &r)) < best) { best = d; *bb = 0; *bc = r; *ba = 1 - *bb - *bc; }
    if ((dg || wbaia < 0) && (d = ve_d2(xi, xa, xb, &r)) < best) { best = d; *bb = r; *bc = 0; *ba = 1 - *bb - *bc; }
  }
#pragma unroll
  for (int k = 0; k < 3; k++) dir->v[k] = xi.v[k] - xa.v[k] * *ba - xb.v[k] * *bb - xc.v[k] * *bc;
  return dot(*dir, *dir);
}

__device__ inline double mn2(double a, double b) { return a < b ? a : b; }
__device__ inline double mx2(double a, double b) { return a > b ? a : b; }
__device__ inline double mn3(double a, double b, double c) { return a < b ? (a < c ? a : c) : (b < c ? b : c); }
__device__ inline double mx3(double a, double b, double c) { return a > b ? (a > c ? a : c) : (b > c ? b : c); }

__device__ inline bool pt_ccd_filter(const P3* x, const P3* dx, double dist) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double tmin = mn3(x[1].v[k], x[2].v[k], x[3].v[k]);
    tmin = mn3(tmin, x[1].v[k] + dx[1].v[k], x[2].v[k] + dx[2].v[k]);
    tmin = mn2(tmin, x[3].v[k] + dx[3].v[k]);
    double tmax = mx3(x[1].v[k], x[2].v[k], x[3].v[k]);
    tmax = mx3(tmax, x[1].v[k] + dx[1].v[k], x[2].v[k] + dx[2].v[k]);
    tmax = mx2(tmax, x[3].v[k] + dx[3].v[k]);
    double p0 = x[0].v[k], p1 = x[0].v[k] + dx[0].v[k];
    double pmin = mn2(p0, p1), pmax = mx2(p0, p1);
    if ((pmin - tmax) > dist || (tmin - pmax) > dist) return false;
  }
  return true;
}
__device__ inline bool ee_ccd_filter(const P3* x, const P3* dx, double dist) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double mina = mn2(x[0].v[k], x[0].v[k] + dx[0].v[k]);
    mina = mn2(mina, x[1].v[k]);
    mina = mn2(mina, x[1].v[k] + dx[1].v[k]);
    double maxa = mx2(x[0].v[k], x[0].v[k] + dx[0].v[k]);
    maxa = mx2(maxa, x[1].v[k]);
    maxa = mx2(maxa, x[1].v[k] + dx[1].v[k]);
    double minb = mn2(x[2].v[k], x[2].v[k] + dx[2].v[k]);
    minb = mn2(minb, x[3].v[k]);
    minb = mn2(minb, x[3].v[k] + dx[3].v[k]);
    double maxb = mx2(x[2].v[k], x[2].v[k] + dx[2].v[k]);
    maxb = mx2(maxb, x[3].v[k]);
    maxb = mx2(maxb, x[3].v[k] + dx[3].v[k]);
    if ((mina - maxb) > dist || (minb - maxa) > dist) return false;
  }
  return true;
}

__device__ inline double ccd_dfun(int type, const P3* x, double* d2out) {
  double a, b, c;
  P3 dir;
  if (type == 0) {
    double d2 = pt_d2(x[0], x[1], x[2], x[3], &a, &b, &c, &dir);
    *d2out = d2;
    return d2;
  }
  double d2 = ee_d2(x[0], x[1], x[2], x[3], &a, &b, &dir);
  if (d2 <= 0) {
    P3 e = sub(x[0], x[2]);
    d2 = dot(e, e);
    e = sub(x[0], x[3]); d2 = mn2(dot(e, e), d2);
    e = sub(x[1], x[2]); d2 = mn2(dot(e, e), d2);
    e = sub(x[1], x[3]); d2 = mn2(dot(e, e), d2);
  }
  *d2out = d2;
  return d2;
}

// ACCD with the reference's constants as parameters
__device__ inline double accd(int type, P3* x, P3* dx, int iters, double step, double gapf, int touchPolicy) {
  P3 mov;
#pragma unroll
  for (int k = 0; k < 3; k++) mov.v[k] = (dx[0].v[k] + dx[1].v[k] + dx[2].v[k] + dx[3].v[k]) / 4.0;
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) dx[i].v[k] -= mov.v[k];
  double maxdisp;
  if (type == 0) {
    double m2 = mx2(dot(dx[1], dx[1]), dot(dx[2], dx[2]));
    m2 = mx2(m2, dot(dx[3], dx[3]));
    maxdisp = sqrt(dot(dx[0], dx[0])) + sqrt(m2);
  } else {
    double m0 = mx2(dot(dx[0], dx[0]), dot(dx[1], dx[1]));
    double m1 = mx2(dot(dx[2], dx[2]), dot(dx[3], dx[3]));
    maxdisp = sqrt(m0) + sqrt(m1);
  }
  if (maxdisp == 0.0) return 1.0;
  double d2;
  double dF = ccd_dfun(type, x, &d2);
  double dcur = sqrt(d2);
  if (touchPolicy == 1 && !(dcur > 0)) return 0.0;
  double gap = gapf * dF / dcur;
  double toc = 0.0;
  for (int i = 0; i < iters; i++) {
    double tlb = step * dF / (dcur * maxdisp);
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int k = 0; k < 3; k++) x[j].v[k] += tlb * dx[j].v[k];
    dF = ccd_dfun(type, x, &d2);
    dcur = sqrt(d2);
    if (touchPolicy == 1 && !(dcur > 0)) return toc;
    if (toc != 0.0 && (dF / dcur) < gap) break;
    toc += tlb;
    if (toc > 1.0) { toc = 1.0; break; }
  }
  return toc;
}

__device__ inline P3 load3(const double* X, int ld, int i) { return {{X[i], X[ld + i], X[2 * ld + i]}}; }

__global__ void k_toi_init(unsigned long long* bits) { *bits = (unsigned long long)__double_as_longlong(1.0); }

__global__ void __launch_bounds__(kBlock) k_ccd(int nRecs, int ld, const int4* __restrict__ recs,
                                                const double* __restrict__ X0, const double* __restrict__ X1,
                                                double filter, int iters, double step, double gapf, int touchPolicy,
                                                unsigned long long* __restrict__ toiBits) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  double toi = 1.0;
  if (t < nRecs) {
    int4 r = recs[t];
    int type = r.x < 0 ? 1 : 0;
    int id[4] = {type ? ~r.x : r.x, r.y, r.z, r.w};
    bool skip = type == 0 ? (id[0] == id[1] || id[0] == id[2] || id[0] == id[3])
                          : (id[0] == id[2] || id[0] == id[3] || id[1] == id[2] || id[1] == id[3]);
    if (!skip) {
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
        if (v < toi) toi = v;  // NaN (reference rule) never wins
      }
    }
  }
  for (int o = kWave / 2; o > 0; o >>= 1) {
    double other = __shfl_xor(toi, o, kWave);
    toi = other < toi ? other : toi;
  }
  if ((threadIdx.x % kWave) == 0 && toi < 1.0) atomicMin(toiBits, (unsigned long long)__double_as_longlong(toi));
}

// barrier (PdIpcGpuFunc.cu:2745-2754)
__device__ inline double barrier(double d2, double dh2, double kappa) {
  double e = 0.0;
  if (d2 < dh2) {
    double t2 = d2 - dh2;
    e = -kappa * t2 * t2 * log(d2 / dh2) / (dh2 * dh2);
  }
  return e;
}

__global__ void __launch_bounds__(kBlock) k_collision(int nRecs, int ld, const int4* __restrict__ recs,
                                                      const double* __restrict__ X, const double* __restrict__ cm,
                                                      double dHat, double kappa, double dt, double* __restrict__ diag,
                                                      double* __restrict__ rhs, int* __restrict__ activeCount) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  bool active = false;
  if (t < nRecs) {
    int4 r = recs[t];
    int type = r.x < 0 ? 1 : 0;
    int id[4] = {type ? ~r.x : r.x, r.y, r.z, r.w};
    bool skip = type == 0 ? (id[0] == id[1] || id[0] == id[2] || id[0] == id[3])
                          : (id[0] == id[2] || id[0] == id[3] || id[1] == id[2] || id[1] == id[3]);
    if (!skip) {
      P3 x[4];
#pragma unroll
      for (int q = 0; q < 4; q++) x[q] = load3(X, ld, id[q]);
      bool pass = true;
      if (type == 0) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
          double tmin = mn3(x[1].v[k], x[2].v[k], x[3].v[k]), tmax = mx3(x[1].v[k], x[2].v[k], x[3].v[k]);
          if ((x[0].v[k] - tmax) > dHat || (tmin - x[0].v[k]) > dHat) pass = false;
        }
      } else {
#pragma unroll
        for (int k = 0; k < 3; k++) {
          double mina = mn2(x[0].v[k], x[1].v[k]), maxa = mx2(x[0].v[k], x[1].v[k]);
          double minb = mn2(x[2].v[k], x[3].v[k]), maxb = mx2(x[2].v[k], x[3].v[k]);
          if ((mina - maxb) > dHat || (minb - maxa) > dHat) pass = false;
        }
      }
      if (pass) {
        const double dh2 = dHat * dHat;
        P3 dir;
        double d2, mp, mq;
        double sgn[4];
        if (type == 0) {
          double a, b, c;
          d2 = pt_d2(x[0], x[1], x[2], x[3], &a, &b, &c, &dir);
          mp = cm[id[0]];
          mq = a * cm[id[1]] + b * cm[id[2]] + c * cm[id[3]];
          sgn[0] = 1; sgn[1] = -1; sgn[2] = -1; sgn[3] = -1;
        } else {
          double rr, ss;
          d2 = ee_d2(x[0], x[1], x[2], x[3], &rr, &ss, &dir);
          mp = (1 - rr) * cm[id[0]] + rr * cm[id[1]];
          mq = (1 - ss) * cm[id[2]] + ss * cm[id[3]];
          sgn[0] = 1; sgn[1] = 1; sgn[2] = -1; sgn[3] = -1;
        }
        if (d2 < dh2) {
          active = true;
          double dist = sqrt(d2);
          double w = barrier(d2, dh2, kappa) + mn2(mp, mq) / (dt * dt);
          double rp = mq / (mp + mq), rq = mp / (mp + mq);
          double len = sqrt(dot(dir, dir));
#pragma unroll
          for (int k = 0; k < 3; k++) dir.v[k] /= len;
          double gapd = dHat - dist;
          double dp = gapd * rp, dq = gapd * rq;
#pragma unroll
          for (int q = 0; q < 4; q++) {
            atomicAdd(diag + id[q], w);
            double sh = sgn[q] > 0 ? dp : -dq;
#pragma unroll
            for (int k = 0; k < 3; k++) atomicAdd(rhs + k * ld + id[q], w * (x[q].v[k] + sh * dir.v[k]));
          }
        }
      }
    }
  }
  unsigned long long b = __ballot(active);
  if ((threadIdx.x % kWave) == 0 && b) atomicAdd(activeCount, (int)__popcll(b));
}

}  // namespace

void launch_toi_init(const DeviceMesh& m, hipStream_t s) {
  hipLaunchKernelGGL(k_toi_init, dim3(1), dim3(1), 0, s, m.toiBits);
}

void launch_ccd(const DeviceMesh& m, int nRecs, double filter, int iters, double step, double gap, int touchPolicy,
                hipStream_t s) {
  if (nRecs <= 0) return;
  hipLaunchKernelGGL(k_ccd, dim3((nRecs + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nRecs, m.ld, m.recs, m.Xold, m.X,
                     filter, iters, step, gap, touchPolicy, m.toiBits);
}

void launch_collision(const DeviceMesh& m, int nRecs, double dhat, double kappa, double dt, hipStream_t s) {
  if (nRecs <= 0) return;
  hipLaunchKernelGGL(k_collision, dim3((nRecs + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nRecs, m.ld, m.recs, m.X,
                     m.colMass, dhat, kappa, dt, m.colDiag, m.colRhs, m.counters + 2);
}

}  // namespace pdipc
