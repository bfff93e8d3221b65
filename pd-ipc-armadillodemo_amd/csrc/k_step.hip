// Per-point kernels of the step driver: prediction, system diagonal, search
// direction, TOI clamp with the residual reduction, velocity update.
// HBM-bound streaming kernels, one lane per point, SoA loads (coalesced).
#include <algorithm>

#include "device.hpp"

namespace pdipc {

namespace {
constexpr int kBlock = 256;
inline int grid_for(int64_t n, int block = kBlock) { return (int)((n + block - 1) / block); }
}  // namespace

// predictiveKernel (Simulator/PD-IPC/PdIpcGpuFunc.cu:88-147): s = x + dt*(v + dt*f/m),
// inertial diagonal m/dt^2 and rhs0 = (m/dt^2) s; fixed points keep x and get v = 0.
__global__ void k_predict(int n, int ld, double dt, double* __restrict__ X, double* __restrict__ V,
                          double* __restrict__ initRhs, double* __restrict__ inertialDiag,
                          const double* __restrict__ gravY, const double* __restrict__ invMass,
                          const int* __restrict__ fixFlag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (fixFlag[i]) {
    V[i] = 0.0; V[ld + i] = 0.0; V[2 * ld + i] = 0.0;
    return;
  }
  double im = invMass[i];
  double m = im * dt;
  double f[3] = {0.0, gravY[i], 0.0};
  double s[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    double v = V[c * ld + i];
    v += m * f[c];
    v *= dt;
    s[c] = v + X[c * ld + i];
  }
  double w = 1.0 / (dt * m);
  inertialDiag[i] = 1.0 / (im * dt * dt);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    initRhs[c * ld + i] = w * s[c];
    X[c * ld + i] = s[c];
  }
}

void launch_predict(const DeviceMesh& m, double dt, hipStream_t s) {
  if (m.nSim == 0) return;
  hipLaunchKernelGGL(k_predict, dim3(grid_for(m.nSim)), dim3(kBlock), 0, s, m.nSim, m.ld, dt, m.X, m.V, m.initRhs,
                     m.inertialDiag, m.gravY, m.invMass, m.fixFlag);
}

// LocalGlobalIteration diagonal (PdIpcSimulator.cpp:1557-1560): D = elastic + inertial
// (+ collision).  Also records which rows changed since the last operator update
// (or were marked stale by the host), so the M^2 update can skip untouched rows.
__global__ void k_diag(int n, const double* __restrict__ e, const double* __restrict__ in,
                       const double* __restrict__ col, int withCol, double* __restrict__ D,
                       uint8_t* __restrict__ stale, uint8_t* __restrict__ changed) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = e[i] + in[i];
  if (withCol) d += col[i];
  uint8_t c = stale[i] | (uint8_t)(d != D[i]);
  changed[i] = c;
  stale[i] = 0;
  D[i] = d;
}

void launch_diag(const DeviceMesh& m, int withCollision, hipStream_t s) {
  hipLaunchKernelGGL(k_diag, dim3(grid_for(m.nSim)), dim3(kBlock), 0, s, m.nSim, m.elasticDiag, m.inertialDiag,
                     m.colDiag, withCollision, m.sysDiag, m.dStale, m.dChanged);
}

// computeDirectionKernel (PdIpcGpuFunc.cu:185-200); thread 0 also resets the
// device-resident TOI minimum to 1 for the CCD pass that follows (a launch
// fewer per outer iteration)
__global__ void k_direction(int n, int ld, const double* __restrict__ X, const double* __restrict__ Xold,
                            double* __restrict__ D, unsigned long long* __restrict__ toiBits) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *toiBits = (unsigned long long)__double_as_longlong(1.0);
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 3; c++) D[c * ld + i] = X[c * ld + i] - Xold[c * ld + i];
}

void launch_direction(const DeviceMesh& m, hipStream_t s) {
  hipLaunchKernelGGL(k_direction, dim3(grid_for(std::max(1, m.nSim))), dim3(kBlock), 0, s, m.nSim, m.ld, m.X, m.Xold,
                     m.Dir, m.toiBits);
}

// clampDirectionKernel (PdIpcGpuFunc.cu:222-238) reading the device-resident TOI,
// fused with the block partial sums of |d|^2 (the residual of PdIpcSimulator.cpp:1774).
// The TOI minimum over all records, scaled by toi_scale when below 1, equals the
// reference's per-block scaling followed by the global minimum (the scaling is monotone).
__global__ void k_clamp_residual(int n, int ld, double toiScale, const unsigned long long* __restrict__ toiBits,
                                 const double* __restrict__ Xold, const double* __restrict__ D, double* __restrict__ X,
                                 double* __restrict__ partials) {
  __shared__ double red[kBlock / kWave];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double tmin = __longlong_as_double((long long)*toiBits);
  double toi = tmin < 1.0 ? tmin * toiScale : 1.0;
  double acc = 0.0;
  if (i < n) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double d = D[c * ld + i];
      X[c * ld + i] = Xold[c * ld + i] + toi * d;
      acc += d * d;
    }
  }
  for (int off = kWave / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, kWave);
  int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
  if (lane == 0) red[wv] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / kWave; w++) t += red[w];
    partials[blockIdx.x] = t;
  }
}

__global__ void k_sum_partials(int n, const double* __restrict__ partials, double* __restrict__ out) {
  __shared__ double red[kBlock / kWave];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partials[i];
  for (int off = kWave / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, kWave);
  int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
  if (lane == 0) red[wv] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / kWave; w++) t += red[w];
    *out = t;
  }
}

void launch_clamp_residual(const DeviceMesh& m, double toiScale, hipStream_t s) {
  int g = grid_for(m.nSim);
  hipLaunchKernelGGL(k_clamp_residual, dim3(g), dim3(kBlock), 0, s, m.nSim, m.ld, toiScale, m.toiBits, m.Xold, m.Dir,
                     m.X, m.partials);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, s, g, m.partials, m.residual);
}

// updateVelocityKernel (PdIpcGpuFunc.cu:262-280) with damping 1 - dt (PdIpcSimulator.cpp:1789)
__global__ void k_velocity(int n, int ld, double invt, const double* __restrict__ X, const double* __restrict__ X0,
                           double* __restrict__ V) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 3; c++) V[c * ld + i] = (X[c * ld + i] - X0[c * ld + i]) * invt;
}

void launch_velocity(const DeviceMesh& m, double dt, hipStream_t s) {
  double invt = 1.0 / dt * (1. - dt);
  hipLaunchKernelGGL(k_velocity, dim3(grid_for(m.nSim)), dim3(kBlock), 0, s, m.nSim, m.ld, invt, m.X, m.X0, m.V);
}

__global__ void k_zero(double* __restrict__ p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}
__global__ void k_copy(double* __restrict__ d, const double* __restrict__ s, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = s[i];
}

// ---- adjoint helpers (simulated rows of [3][ld] vectors)
__global__ void k_axpby(int n, int ld, double a, const double* __restrict__ x, double b, const double* __restrict__ y,
                        double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    double v = a * x[c * ld + i];
    if (y) v += b * y[c * ld + i];
    out[c * ld + i] = v;
  }
}
void launch_axpby(const DeviceMesh& m, double a, const double* x, double b, const double* y, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(m.nSim)), dim3(kBlock), 0, s, m.nSim, m.ld, a, x, b, y, out);
}

__global__ void k_div_rows(int n, int ld, const double* __restrict__ x, const double* __restrict__ D,
                           double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = D[i];
#pragma unroll
  for (int c = 0; c < 3; c++) out[c * ld + i] = x[c * ld + i] / d;
}
void launch_div_rows(const DeviceMesh& m, const double* x, const double* D, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_div_rows, dim3(grid_for(m.nSim)), dim3(kBlock), 0, s, m.nSim, m.ld, x, D, out);
}

__global__ void k_mul_rows(int n, int ld, const double* __restrict__ x, const double* __restrict__ D,
                           double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = D[i];
#pragma unroll
  for (int c = 0; c < 3; c++) out[c * ld + i] = x[c * ld + i] * d;
}
void launch_mul_rows(const DeviceMesh& m, const double* x, const double* D, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_mul_rows, dim3(grid_for(m.nSim)), dim3(kBlock), 0, s, m.nSim, m.ld, x, D, out);
}

// scale * sum_i a_i . (x1_i - x0_i): the adjoint of the TOI (deterministic two-pass sum)
__global__ void k_dot_dir(int n, int ld, const double* __restrict__ a, const double* __restrict__ x1,
                          const double* __restrict__ x0, double* __restrict__ partials) {
  __shared__ double red[kBlock / kWave];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (i < n) {
#pragma unroll
    for (int c = 0; c < 3; c++) acc += a[c * ld + i] * (x1[c * ld + i] - x0[c * ld + i]);
  }
  for (int off = kWave / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, kWave);
  int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
  if (lane == 0) red[wv] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / kWave; w++) t += red[w];
    partials[blockIdx.x] = t;
  }
}
__global__ void k_scale_scalar(double* p, double scale) { *p *= scale; }
void launch_dot_dir(const DeviceMesh& m, const double* a, const double* x1, const double* x0, double scale,
                    double* partials, double* out, hipStream_t s) {
  int g = grid_for(m.nSim);
  hipLaunchKernelGGL(k_dot_dir, dim3(g), dim3(kBlock), 0, s, m.nSim, m.ld, a, x1, x0, partials);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, s, g, partials, out);
  hipLaunchKernelGGL(k_scale_scalar, dim3(1), dim3(1), 0, s, out, scale);
}

// adjoint of the prediction s = x + dt (v + dt f / m), rhs0 = (m / dt^2) s, and of
// x0 = x:  xbar = axn + sbar, vbar = dt sbar with sbar = as + (m / dt^2) rhs0bar
// (fixed points keep x, so sbar = as and vbar = 0)
__global__ void k_predict_adj(int n, int ld, double dt, const int* __restrict__ fixFlag,
                              const double* __restrict__ inertialDiag, const double* __restrict__ as,
                              const double* __restrict__ rhs0bar, const double* __restrict__ axn,
                              double* __restrict__ xbar, double* __restrict__ vbar) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool fx = fixFlag[i] != 0;
  double w = inertialDiag[i];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    int idx = c * ld + i;
    double sb = fx ? as[idx] : as[idx] + w * rhs0bar[idx];
    xbar[idx] = axn[idx] + sb;
    vbar[idx] = fx ? 0.0 : dt * sb;
  }
}
void launch_predict_adj(const DeviceMesh& m, double dt, const double* as, const double* rhs0bar, const double* axn,
                        double* xbar, double* vbar, hipStream_t s) {
  hipLaunchKernelGGL(k_predict_adj, dim3(grid_for(m.nSim)), dim3(kBlock), 0, s, m.nSim, m.ld, dt, m.fixFlag,
                     m.inertialDiag, as, rhs0bar, axn, xbar, vbar);
}

__global__ void k_add(double* __restrict__ d, const double* __restrict__ s, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] += s[i];
}
void launch_add_rows(double* dst, const double* src, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_add, dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, n);
}

void launch_zero(double* p, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_zero, dim3(grid_for(n)), dim3(kBlock), 0, s, p, n);
}
void launch_copy(double* dst, const double* src, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, n);
}

}  // namespace pdipc
