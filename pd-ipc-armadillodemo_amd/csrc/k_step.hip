// Per-point kernels of the step driver: prediction, system diagonal, search
// direction, TOI clamp with the residual reduction, velocity update.
// HBM-bound streaming kernels, one lane per point, SoA loads (coalesced).
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

// computeDirectionKernel (PdIpcGpuFunc.cu:185-200)
__global__ void k_direction(int n, int ld, const double* __restrict__ X, const double* __restrict__ Xold,
                            double* __restrict__ D) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 3; c++) D[c * ld + i] = X[c * ld + i] - Xold[c * ld + i];
}

void launch_direction(const DeviceMesh& m, hipStream_t s) {
  hipLaunchKernelGGL(k_direction, dim3(grid_for(m.nSim)), dim3(kBlock), 0, s, m.nSim, m.ld, m.X, m.Xold, m.Dir);
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

void launch_zero(double* p, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_zero, dim3(grid_for(n)), dim3(kBlock), 0, s, p, n);
}
void launch_copy(double* dst, const double* src, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, n);
}

}  // namespace pdipc
