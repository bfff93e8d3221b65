// The global step: Chebyshev-accelerated two-step weighted Jacobi on the fixed
// PD system (chebyshevJacobiR2SolverHost, Simulator/PD-IPC/PdIpcGpuFunc.cu:1245-1346).
//
//   M^2 = ((1-w) I + w D^{-1} R)^2         (the "R2" row operators, :899-974)
//   c   = w^2 D^{-1} R D^{-1} b + (2-w) w D^{-1} b          (:1041-1074)
//   x1  = M^2 x0 + c                                        (:1076-1129)
//   x_{k+1} = W_k (M^2 x_k + c - x_{k-1}) + x_{k-1}         (:1185-1242)
//   with r = rho^2, W_1 = 2/(2-r), W_{k+1} = 4/(4 - r W_k)  (:1318-1340)
//
// M^2 is stored in a SELL layout (slices of 64/G rows, G lanes per row, one
// wavefront per slice, each lane's entries consecutive), so the matrix stream
// -- the dominant HBM traffic of the whole step -- is read in contiguous
// wavefront-wide blocks (16-byte loads per lane at the common width 8).  Each
// matrix row is read once for all three coordinates (the reference re-reads it
// per coordinate).  When all slices have one width uw (set-up pads them when
// that costs at most 2% more slots) a sweep computes its slice's offset
// instead of loading it; otherwise rows are sorted by width within windows of
// two slices and a slot's row is read from perm.  Columns are 16-bit offsets
// from the entry's slot (its row, unless reordered) whenever every one fits, so
// the gathers never wait for perm.  The operator update rebuilds the M^2
// values from the 2-path lists once per outer iteration.
//
// FP32-split: the sweeps run in FP32 on the correction delta = x - x0, driven
// by the FP64 residual of the fixed-point map r0 = M^2 x0 + c - x0 (computed
// as M (M x0) + c - x0 over R1, two light passes instead of the M^2 pattern):
//   delta_1 = r0,  delta_{k+1} = W_k (M^2 delta_k + r0 - delta_{k-1}) + delta_{k-1}
// which is algebraically the same iterate; FP32 only rounds the (small)
// correction, never the absolute coordinates.
#include <type_traits>

#include "device.hpp"
#include "host_setup.hpp"  // kTileSlices

namespace pdipc {
namespace {

constexpr int kBlock = 256;

// M^2_ij = (w^2 sum_k R1_ik R1_kj / D_k + 2 w (1 - w) R1_ij) / D_i + (1 - w)^2 [i == j]
// (updateR1/R2/R2RowOperatorsKernel, PdIpcGpuFunc.cu:899-974), recomputed only
// for rows whose value can have changed: a row depends on D_i and on D_k of its
// R1 neighbours k, so rows with no changed diagonal within one R1 hop keep
// their (bit-identical) values.
__global__ void k_dirty_rows(int n, const int* __restrict__ off, const int* __restrict__ col,
                             const uint8_t* __restrict__ changed, uint8_t* __restrict__ dirty) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint8_t d = changed[i];
  for (int q = off[i]; q < off[i + 1] && !d; q++) d |= changed[col[q]];
  dirty[i] = d;
}

__global__ void k_operators(int64_t nE, double w, const int* __restrict__ rowFlag, const uint8_t* __restrict__ dirty,
                            const double* __restrict__ r1, const int64_t* __restrict__ pOff,
                            const int* __restrict__ pK, const double* __restrict__ pA,
                            const double* __restrict__ D, double* __restrict__ val, float* __restrict__ valF) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nE) return;
  int rf = rowFlag[e];
  int row = rf & 0x7fffffff;
  if (!dirty[row]) return;
  double s = 0.0;
  for (int64_t p = pOff[e]; p < pOff[e + 1]; p++) s += pA[p] / D[pK[p]];
  double v = (w * w * s + 2.0 * (1.0 - w) * w * r1[e]) / D[row];
  if (rf < 0) v += (1.0 - w) * (1.0 - w);
  val[e] = v;
  valF[e] = (float)v;
}

template <int G, class V>
__device__ inline void group_reduce(V acc[3]) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
    for (int k = 0; k < 3; k++) acc[k] += __shfl_xor(acc[k], o, kWave);
}

struct FirstTermArgs {  // c = (w^2 sum_j R1_ij b_j / D_j + (2 - w) w b_i) / D_i   (PdIpcGpuFunc.cu:1041-1074)
  const int* off;
  const int* col;
  const double* val;
  const double* D;
  const double* b;
  double w;
};

// The first FP64 sweep of a solve, reading x0 = X (SoA) and computing c on the
// fly (the first term is fused: the G lanes of a row split its R1 entries):
// c -> cA, x0 -> xa0, x1 = M^2 x0 + c -> xa1 (and out, SoA, when given).
template <int G, bool C16>
// (at most 64 registers: eight wavefronts per SIMD, so a 65 k-row sweep runs in one round)
__global__ void __launch_bounds__(kBlock, 8) k_first64(int n, int nSlices, int ld,
                                                       const int64_t* __restrict__ sliceOff,
                                                       const int* __restrict__ sliceW, int uw, const int* __restrict__ perm,
                                                       const int* __restrict__ col, const int16_t* __restrict__ col16,
                                                       const double* __restrict__ val, const double* __restrict__ X,
                                                       Row64* __restrict__ cA, Row64* __restrict__ xa0,
                                                       Row64* __restrict__ xa1, double* __restrict__ out,
                                                       FirstTermArgs ft) {
  int gt = blockIdx.x * blockDim.x + threadIdx.x;
  int slice = gt / kWave, lane = gt % kWave;
  if (slice >= nSlices) return;  // whole wavefront
  constexpr int S = kWave / G;
  const int slot = slice * S + lane / G;
  const int row = perm ? perm[slot] : slot;
  int sub = lane % G;
  const int width = uw > 0 ? uw : sliceW[slice];
  const bool lm = uw == kSellLaneMajorW;  // lane-major slices, else chunk-major
  const int64_t off = (uw > 0 ? (int64_t)slice * uw * kWave : sliceOff[slice]) + (lm ? (int64_t)lane * width : lane);
  const int step = lm ? 1 : kWave;
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll 4
  for (int c = 0; c < width; c++) {
    int64_t e = off + (int64_t)c * step;
    int j = C16 ? slot + (int)col16[e] : col[e];
    double v = val[e];
    acc[0] += v * X[j];
    acc[1] += v * X[ld + j];
    acc[2] += v * X[2 * ld + j];
  }
  group_reduce<G>(acc);
  double cr[3] = {0.0, 0.0, 0.0};
  if (row < n)
    for (int q = ft.off[row] + sub; q < ft.off[row + 1]; q += G) {
      int j = ft.col[q];
      double f = ft.val[q] / ft.D[j];
#pragma unroll
      for (int k = 0; k < 3; k++) cr[k] += f * ft.b[k * ld + j];
    }
  group_reduce<G>(cr);
  if (sub != 0 || row >= n) return;
  double di = ft.D[row];
  double c[3], y[3], x0[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int idx = k * ld + row;
    c[k] = (ft.w * ft.w * cr[k] + (2.0 - ft.w) * ft.w * ft.b[idx]) / di;
    y[k] = acc[k] + c[k];
    x0[k] = X[idx];
  }
  cA[row] = Row64{c[0], c[1], c[2], 0.0};
  xa0[row] = Row64{x0[0], x0[1], x0[2], 0.0};
  xa1[row] = Row64{y[0], y[1], y[2], 0.0};
  if (out) {
    out[row] = y[0];
    out[ld + row] = y[1];
    out[2 * ld + row] = y[2];
  }
}

// FP32-split start: the FP64 residual r0 = M^2 x + c - x of x = X, with M^2 x
// applied as M (M x) (M = (1 - w) I + w D^{-1} R1: two passes over the ~7x
// smaller R1 instead of one over the M^2 pattern; equal up to rounding), then
// delta_0 = 0, delta_1 = r0.  With y = M x and u = D^{-1} b,
//   M^2 x + c - x = (1 - w) y + (2 - w) w u - x  +  w D^{-1} R1 (y + w u),
// so pass A gathers x only and writes v = y + w u and the pointwise part pt;
// pass B adds w D^{-1} R1 v.  G lanes per R1 row.
template <int G>
__global__ void __launch_bounds__(kBlock) k_res_a(int n, int ld, double w, const int* __restrict__ off,
                                                  const int* __restrict__ col, const double* __restrict__ val,
                                                  const double* __restrict__ D, const double* __restrict__ X,
                                                  const double* __restrict__ b, Row64* __restrict__ v,
                                                  Row64* __restrict__ pt) {
  int gt = blockIdx.x * blockDim.x + threadIdx.x;
  int row = gt / G, sub = gt % G;
  double sx[3] = {0.0, 0.0, 0.0};
  if (row < n)
    for (int q = off[row] + sub; q < off[row + 1]; q += G) {
      int j = col[q];
      double a = val[q];
      sx[0] += a * X[j];
      sx[1] += a * X[ld + j];
      sx[2] += a * X[2 * ld + j];
    }
  group_reduce<G>(sx);
  if (sub != 0 || row >= n) return;
  double di = D[row];
  double vv[3], pp[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double x = X[k * ld + row];
    double y = (1.0 - w) * x + w * sx[k] / di;
    double u = b[k * ld + row] / di;
    vv[k] = y + w * u;
    pp[k] = (1.0 - w) * y + (2.0 - w) * w * u - x;
  }
  v[row] = Row64{vv[0], vv[1], vv[2], 0.0};
  pt[row] = Row64{pp[0], pp[1], pp[2], 0.0};
}

template <int G>
__global__ void __launch_bounds__(kBlock) k_res_b(int n, double w, const int* __restrict__ off,
                                                  const int* __restrict__ col, const double* __restrict__ val,
                                                  const double* __restrict__ D, const Row64* __restrict__ v,
                                                  const Row64* __restrict__ pt, float4* __restrict__ dpreF,
                                                  float4* __restrict__ dnextF, float4* __restrict__ r0F) {
  int gt = blockIdx.x * blockDim.x + threadIdx.x;
  int row = gt / G, sub = gt % G;
  // the row's own operands first, so that they arrive with the gathers
  const int rr = row < n ? row : n - 1;
  const double dr = D[rr];
  const Row64 p = pt[rr];
  double sv[3] = {0.0, 0.0, 0.0};
  if (row < n)
    for (int q = off[row] + sub; q < off[row + 1]; q += G) {
      double a = val[q];
      Row64 vj = v[col[q]];
      sv[0] += a * vj.x;
      sv[1] += a * vj.y;
      sv[2] += a * vj.z;
    }
  group_reduce<G>(sv);
  if (sub != 0 || row >= n) return;
  double f = w / dr;
  float4 rv = make_float4((float)(p.x + f * sv[0]), (float)(p.y + f * sv[1]), (float)(p.z + f * sv[2]), 0.0f);
  r0F[row] = rv;
  dpreF[row] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  dnextF[row] = rv;
}

// FP64 Chebyshev step x_{k+1} = W (M^2 x_k + c - x_{k-1}) + x_{k-1} on Row64
// vectors (one 32-byte gather per matrix entry, U entries of a lane fetched
// together); out (SoA), when given, receives x_{k+1} as well.
template <int G, bool C16>
__global__ void __launch_bounds__(kBlock) k_sweep64(int n, int nSlices, int ld, double W,
                                                    const int64_t* __restrict__ sliceOff,
                                                    const int* __restrict__ sliceW, int uw, const int* __restrict__ perm, const int* __restrict__ col,
                                                    const int16_t* __restrict__ col16, const double* __restrict__ val,
                                                    const Row64* __restrict__ cur, const Row64* __restrict__ pre,
                                                    const Row64* __restrict__ cA,
                                                    Row64* __restrict__ next, double* __restrict__ out) {
  constexpr int U = 4;
  int gt = blockIdx.x * blockDim.x + threadIdx.x;
  int slice = gt / kWave, lane = gt % kWave;
  if (slice >= nSlices) return;
  constexpr int S = kWave / G;
  const int slot = slice * S + lane / G;  // < ld: slices cover at most ld rows
  const int row = perm ? perm[slot] : slot;
  const int width = uw > 0 ? uw : sliceW[slice];
  const bool lm = uw == kSellLaneMajorW;  // lane-major slices, else chunk-major
  const int64_t off = (uw > 0 ? (int64_t)slice * uw * kWave : sliceOff[slice]) + (lm ? (int64_t)lane * width : lane);
  const int step = lm ? 1 : kWave;
  Row64 p = pre[row], c = cA[row];  // loaded up front, with the matrix
  double acc[3] = {0.0, 0.0, 0.0};
  for (int c0 = 0; c0 < width; c0 += U) {
    int j[U];
    double v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      j[u] = row;
      v[u] = 0.0;
      if (c0 + u < width) {
        int64_t e = off + (int64_t)(c0 + u) * step;
        j[u] = C16 ? slot + (int)col16[e] : col[e];
        v[u] = val[e];
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      Row64 x = cur[j[u]];
      acc[0] += v[u] * x.x;
      acc[1] += v[u] * x.y;
      acc[2] += v[u] * x.z;
    }
  }
  group_reduce<G>(acc);
  if ((lane % G) != 0 || row >= n) return;
  double x0 = W * (acc[0] + c.x - p.x) + p.x;
  double x1 = W * (acc[1] + c.y - p.y) + p.y;
  double x2 = W * (acc[2] + c.z - p.z) + p.z;
  next[row] = Row64{x0, x1, x2, 0.0};
  if (out) {
    out[row] = x0;
    out[ld + row] = x1;
    out[2 * ld + row] = x2;
  }
}

__global__ void k_row64_to_soa(int n, int ld, const Row64* __restrict__ a, double* __restrict__ X) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Row64 v = a[i];
  X[i] = v.x;
  X[ld + i] = v.y;
  X[2 * ld + i] = v.z;
}

// FP32 Chebyshev step on the correction; LAST also writes x = x0 + delta.
// The correction vectors hold one float4 per row, so each matrix entry costs
// one 16-byte gather; the entries of a lane are fetched U at a time (all
// column / value loads first, then all gathers) to keep several memory
// round trips in flight per wavefront.  C16: columns as 16-bit offsets from
// the slot.
template <int G, int LAST, bool C16>
__global__ void __launch_bounds__(kBlock) k_sweep32(int n, int nSlices, int ld, float W,
                                                    const int64_t* __restrict__ sliceOff,
                                                    const int* __restrict__ sliceW, int uw, const int* __restrict__ perm, const int* __restrict__ col,
                                                    const int16_t* __restrict__ col16, const float* __restrict__ val,
                                                    const float4* __restrict__ dcur, const float4* __restrict__ dpre,
                                                    const float4* __restrict__ r0, float4* __restrict__ dnext,
                                                    double* __restrict__ X) {
  constexpr int U = 4;
  int gt = blockIdx.x * blockDim.x + threadIdx.x;
  int slice = gt / kWave, lane = gt % kWave;
  if (slice >= nSlices) return;
  constexpr int S = kWave / G;
  const int slot = slice * S + lane / G;  // < ld: slices cover at most ld rows
  const int row = perm ? perm[slot] : slot;
  const int width = uw > 0 ? uw : sliceW[slice];
  const bool lm = uw == kSellLaneMajorW;  // lane-major slices, else chunk-major
  const int64_t off = (uw > 0 ? (int64_t)slice * uw * kWave : sliceOff[slice]) + (lm ? (int64_t)lane * width : lane);
  const int step = lm ? 1 : kWave;
  // the row's own operands, loaded up front so they arrive with the matrix
  // (the last sweep of a solve also adds into X: its X too)
  float4 p = dpre[row], r = r0[row];
  double xo[3] = {0.0, 0.0, 0.0};
  if (LAST) {
    xo[0] = X[row];
    xo[1] = X[ld + row];
    xo[2] = X[2 * ld + row];
  }
  float acc[3] = {0.f, 0.f, 0.f};
  for (int c0 = 0; c0 < width; c0 += U) {
    int j[U];
    float v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      j[u] = row;
      v[u] = 0.0f;
      if (c0 + u < width) {
        int64_t e = off + (int64_t)(c0 + u) * step;
        j[u] = C16 ? slot + (int)col16[e] : col[e];
        v[u] = val[e];
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      float4 x = dcur[j[u]];
      acc[0] += v[u] * x.x;
      acc[1] += v[u] * x.y;
      acc[2] += v[u] * x.z;
    }
  }
  group_reduce<G>(acc);
  if ((lane % G) != 0 || row >= n) return;
  float d0 = W * (acc[0] + r.x - p.x) + p.x;
  float d1 = W * (acc[1] + r.y - p.y) + p.y;
  float d2 = W * (acc[2] + r.z - p.z) + p.z;
  if (LAST) {
    X[row] = xo[0] + (double)d0;
    X[ld + row] = xo[1] + (double)d1;
    X[2 * ld + row] = xo[2] + (double)d2;
  } else {
    dnext[row] = make_float4(d0, d1, d2, 0.0f);
  }
}

// LDS-staged sweeps.  One workgroup runs a tile of kTileSlices slices: it
// first copies the gathered operand at the tile's column union into LDS (the
// union's columns as 16-bit offsets from the tile's smallest column, listed at
// a fixed stride per tile and padded with its last column, so the list loads
// go out with the kernel's first instructions and every lane stores its rows
// unconditionally), while the lanes' first kChunk matrix entries are already
// on their way; after one barrier every entry reads its operand from LDS (its
// union index `loc`) instead of L2.  LDS holds the staged rows as (x, y) pairs
// and z values (8 + 4 bytes in FP32): reads of whole 16-byte rows at the
// entries' scattered union indices conflicted in the LDS banks more (one
// sweep of a 256^2 cloth: 5.8 -> 4.9 us with the fixed-stride lists).  On a
// mesh-ordered cloth a 32-row tile touches about 9 runs of ~40 columns (the
// 4-ring stencil of M^2), so the L2 traffic of the gathers drops from 16 B
// per entry to about 6 B per row.  Workgroup b runs tile b (the hardware deals
// consecutive workgroups to the 8 XCDs round-robin; giving each XCD a
// contiguous range of tiles, whose unions overlap, measured slower in the
// step: 5.2-5.5 -> 5.6-5.9 us for a 256^2 cloth, 261 -> 279 us at the 3 x 1024^2
// scale).  Entries are summed in the same
// order as the L2-gather kernels, so the results are bit-identical to theirs.
// Entry offsets are 32-bit (staging requires fewer than 2^29 entries): each
// load is a 64-bit base in scalar registers plus a 32-bit lane offset, and a
// lane's loads past its slice's width re-read its last entry (no branches)
// and are weighted 0.  UW = kSellLaneMajorW: every slice has that width and
// is lane-major (known at compile time, so the chunk loop and its guards fold
// away and a lane's entries load as 16-byte vectors); UW = 0: chunk-major
// slices of any width.
constexpr int kChunk32 = 8;  // matrix entries per lane loaded together (FP32 values)
constexpr int kChunk64 = 4;  // (FP64 values)
static_assert(kChunk32 == kSellLaneMajorW, "one FP32 chunk is a lane's whole lane-major run");

template <class T>
__device__ __forceinline__ T ld32(const T* __restrict__ base, uint32_t i) {
  return *(const T*)((const char*)base + i * (uint32_t)sizeof(T));
}

// chunk-major slices: entry c of the lane at off + c * 64
template <int CH, class V>
__device__ __forceinline__ void load_chunk_any(const V* __restrict__ val, const uint16_t* __restrict__ loc,
                                               uint32_t off, int c0, int width, V (&v)[CH], int (&li)[CH]) {
#pragma unroll
  for (int u = 0; u < CH; u++) {
    const int c = c0 + u < width ? c0 + u : width - 1;
    const uint32_t e = off + (uint32_t)(c * kWave);
    const V x = ld32(val, e);
    li[u] = ld32(loc, e);
    v[u] = c0 + u < width ? x : V(0);
  }
}

// Lane-major slices of width 8: a lane's entries are 8 consecutive, 8-aligned slots,
// so a chunk is two 16-byte value loads and one index load (FP32: 8 values,
// one 16-byte index load; FP64: 4 values, one 8-byte index load).
template <int CH, class V>
__device__ __forceinline__ void load_chunk_w8(const V* __restrict__ val, const uint16_t* __restrict__ loc,
                                              uint32_t off, int c0, V (&v)[CH], int (&li)[CH]) {
  const uint32_t e = off + (uint32_t)c0;
  if constexpr (sizeof(V) == 4) {
    static_assert(CH == 8, "FP32 chunks of 8 entries");
    const float4 a = ld32((const float4*)val, e / 4), b = ld32((const float4*)val, e / 4 + 1);
    const uint4 l = ld32((const uint4*)loc, e / 8);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    li[0] = l.x & 0xffff; li[1] = l.x >> 16; li[2] = l.y & 0xffff; li[3] = l.y >> 16;
    li[4] = l.z & 0xffff; li[5] = l.z >> 16; li[6] = l.w & 0xffff; li[7] = l.w >> 16;
  } else if constexpr (CH == 8) {
    const double2 a = ld32((const double2*)val, e / 2), b = ld32((const double2*)val, e / 2 + 1);
    const double2 c = ld32((const double2*)val, e / 2 + 2), d = ld32((const double2*)val, e / 2 + 3);
    const uint4 l = ld32((const uint4*)loc, e / 8);
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    v[4] = c.x; v[5] = c.y; v[6] = d.x; v[7] = d.y;
    li[0] = l.x & 0xffff; li[1] = l.x >> 16; li[2] = l.y & 0xffff; li[3] = l.y >> 16;
    li[4] = l.z & 0xffff; li[5] = l.z >> 16; li[6] = l.w & 0xffff; li[7] = l.w >> 16;
  } else {
    static_assert(CH == 4, "FP64 chunks of 4 or 8 entries");
    const double2 a = ld32((const double2*)val, e / 2), b = ld32((const double2*)val, e / 2 + 1);
    const uint2 l = ld32((const uint2*)loc, e / 4);
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    li[0] = l.x & 0xffff; li[1] = l.x >> 16; li[2] = l.y & 0xffff; li[3] = l.y >> 16;
  }
}

template <int UW, int CH, class V>
__device__ __forceinline__ void load_chunk(const V* __restrict__ val, const uint16_t* __restrict__ loc, uint32_t off,
                                           int c0, int width, V (&v)[CH], int (&li)[CH]) {
  static_assert(UW == 0 || UW == kSellLaneMajorW, "lane-major slices have width kSellLaneMajorW");
  if constexpr (UW == kSellLaneMajorW) load_chunk_w8(val, loc, off, c0, v, li);
  else load_chunk_any(val, loc, off, c0, width, v, li);
}

// LDS rows of a staged tile: (x, y) pairs then z values, SP * kBlock of each
template <class S>
struct StagedRows {
  using S2 = typename std::conditional<sizeof(S) == 4, float2, double2>::type;
  S2* xy;
  S* z;
  __device__ StagedRows(void* lds, int cap) : xy((S2*)lds), z((S*)((S2*)lds + cap)) {}
  __device__ void put(int u, S x, S y, S w) const {
    xy[u] = S2{x, y};
    z[u] = w;
  }
};
template <class S>
constexpr size_t staged_lds_bytes(int sp) { return (size_t)sp * kBlock * 3 * sizeof(S); }

// the tile's union columns, SP rows per lane (lanes past the stride re-read
// the list's last entry)
template <int SP>
__device__ __forceinline__ void tile_columns(int tile, int stride, const int* __restrict__ tileBase,
                                             const uint16_t* __restrict__ tileCol, int (&c)[SP]) {
  const int base = tileBase[tile];
  const uint32_t t0 = (uint32_t)tile * (uint32_t)stride;
#pragma unroll
  for (int q = 0; q < SP; q++) {
    const int u = threadIdx.x + q * kBlock;
    c[q] = base + (int)ld32(tileCol, t0 + (uint32_t)(u < stride ? u : stride - 1));
  }
}

template <int SP, class T>
__device__ __forceinline__ void stage_tile(StagedRows<typename std::conditional<sizeof(T) == 16, float, double>::type> rows,
                                           const T* __restrict__ src, int tile, int stride,
                                           const int* __restrict__ tileBase, const uint16_t* __restrict__ tileCol) {
  int c[SP];
  tile_columns<SP>(tile, stride, tileBase, tileCol, c);
  T v[SP];
#pragma unroll
  for (int q = 0; q < SP; q++) v[q] = ld32(src, (uint32_t)c[q]);
#pragma unroll
  for (int q = 0; q < SP; q++) rows.put(threadIdx.x + q * kBlock, v[q].x, v[q].y, v[q].z);
}

// stage_tile for an SoA operand (x, y, z at src, src + ld, src + 2 ld): a
// staged row is three loads
template <int SP>
__device__ __forceinline__ void stage_tile_soa(StagedRows<double> rows, const double* __restrict__ src, int ld,
                                               int tile, int stride, const int* __restrict__ tileBase,
                                               const uint16_t* __restrict__ tileCol) {
  int c[SP];
  tile_columns<SP>(tile, stride, tileBase, tileCol, c);
  double g[SP][3];
#pragma unroll
  for (int q = 0; q < SP; q++)
#pragma unroll
    for (int k = 0; k < 3; k++) g[q][k] = ld32(src, (uint32_t)(k * ld + c[q]));
#pragma unroll
  for (int q = 0; q < SP; q++) rows.put(threadIdx.x + q * kBlock, g[q][0], g[q][1], g[q][2]);
}

template <int G, int LAST, int SP, int UW>
__global__ void __launch_bounds__(kBlock, 8) k_sweep32s(int n, int nSlices, int ld, float W,
                                                        const int64_t* __restrict__ sliceOff,
                                                        const int* __restrict__ sliceW, int uw,
                                                        const int* __restrict__ perm, const uint16_t* __restrict__ loc,
                                                        const float* __restrict__ val, int stride,
                                                        const int* __restrict__ tileBase,
                                                        const uint16_t* __restrict__ tileCol,
                                                        const float4* __restrict__ dcur, const float4* __restrict__ dpre,
                                                        const float4* __restrict__ r0, float4* __restrict__ dnext,
                                                        double* __restrict__ X) {
  extern __shared__ float2 lds32[];
  const StagedRows<float> xs(lds32, SP * kBlock);
  const int tile = blockIdx.x;
  const int slice = tile * kTileSlices + (int)threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const bool live = slice < nSlices;  // every lane stages and meets the barrier
  constexpr int S = kWave / G;
  const int sl = live ? slice : nSlices - 1;
  const int slot = sl * S + lane / G;
  const int row = perm ? perm[slot] : slot;
  const int width = UW > 0 ? UW : (uw > 0 ? uw : sliceW[sl]);
  const uint32_t off = (uint32_t)(UW > 0 || uw > 0 ? (int64_t)sl * width * kWave : sliceOff[sl]) +
                       (uint32_t)(UW == kSellLaneMajorW ? lane * width : lane);
  float v[kChunk32];
  int li[kChunk32];
  load_chunk<UW>(val, loc, off, 0, width, v, li);
  const float4 p = ld32(dpre, (uint32_t)row), r = ld32(r0, (uint32_t)row);
  stage_tile<SP>(xs, dcur, tile, stride, tileBase, tileCol);
  __syncthreads();
  if (!live) return;
  float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll 1
  for (int c0 = 0;;) {
#pragma unroll
    for (int u = 0; u < kChunk32; u++) {
      const float2 a = xs.xy[li[u]];
      const float z = xs.z[li[u]];
      acc[0] += v[u] * a.x;
      acc[1] += v[u] * a.y;
      acc[2] += v[u] * z;
    }
    c0 += kChunk32;
    if (c0 >= width) break;
    load_chunk<UW>(val, loc, off, c0, width, v, li);
  }
  group_reduce<G>(acc);
  if ((lane % G) != 0 || row >= n) return;
  float d0 = W * (acc[0] + r.x - p.x) + p.x;
  float d1 = W * (acc[1] + r.y - p.y) + p.y;
  float d2 = W * (acc[2] + r.z - p.z) + p.z;
  if (LAST) {
    X[row] += (double)d0;
    X[ld + row] += (double)d1;
    X[2 * ld + row] += (double)d2;
  } else {
    dnext[row] = make_float4(d0, d1, d2, 0.0f);
  }
}

template <int G, int SP, int UW>
__global__ void __launch_bounds__(kBlock, 8) k_sweep64s(int n, int nSlices, int ld, double W,
                                                        const int64_t* __restrict__ sliceOff,
                                                        const int* __restrict__ sliceW, int uw,
                                                        const int* __restrict__ perm, const uint16_t* __restrict__ loc,
                                                        const double* __restrict__ val, int stride,
                                                        const int* __restrict__ tileBase,
                                                        const uint16_t* __restrict__ tileCol,
                                                        const Row64* __restrict__ cur, const Row64* __restrict__ pre,
                                                        const Row64* __restrict__ cA, Row64* __restrict__ next,
                                                        double* __restrict__ out) {
  extern __shared__ double2 lds64[];
  const StagedRows<double> xs(lds64, SP * kBlock);
  const int tile = blockIdx.x;
  const int slice = tile * kTileSlices + (int)threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const bool live = slice < nSlices;
  constexpr int S = kWave / G;
  const int sl = live ? slice : nSlices - 1;
  const int slot = sl * S + lane / G;
  const int row = perm ? perm[slot] : slot;
  const int width = UW > 0 ? UW : (uw > 0 ? uw : sliceW[sl]);
  const uint32_t off = (uint32_t)(UW > 0 || uw > 0 ? (int64_t)sl * width * kWave : sliceOff[sl]) +
                       (uint32_t)(UW == kSellLaneMajorW ? lane * width : lane);
  constexpr int CH = UW == kSellLaneMajorW ? kSellLaneMajorW : kChunk64;  // width 8: the lane's 8 entries in one chunk
  double v[CH];
  int li[CH];
  load_chunk<UW>(val, loc, off, 0, width, v, li);
  stage_tile<SP>(xs, cur, tile, stride, tileBase, tileCol);
  __syncthreads();
  if (!live) return;
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
  for (int c0 = 0;;) {
#pragma unroll
    for (int u = 0; u < CH; u++) {
      const double2 a = xs.xy[li[u]];
      const double z = xs.z[li[u]];
      acc[0] += v[u] * a.x;
      acc[1] += v[u] * a.y;
      acc[2] += v[u] * z;
    }
    c0 += CH;
    if (c0 >= width) break;
    load_chunk<UW>(val, loc, off, c0, width, v, li);
  }
  group_reduce<G>(acc);
  if ((lane % G) != 0 || row >= n) return;
  const Row64 p = ld32(pre, (uint32_t)row), c = ld32(cA, (uint32_t)row);
  double x0 = W * (acc[0] + c.x - p.x) + p.x;
  double x1 = W * (acc[1] + c.y - p.y) + p.y;
  double x2 = W * (acc[2] + c.z - p.z) + p.z;
  next[row] = Row64{x0, x1, x2, 0.0};
  if (out) {
    out[row] = x0;
    out[ld + row] = x1;
    out[2 * ld + row] = x2;
  }
}

// The first FP64 sweep (k_first64) with x0 = X staged in LDS as in
// k_sweep64s; the same sums in the same order, so bit-identical to it.
template <int G, int SP, int UW>
__global__ void __launch_bounds__(kBlock, 8) k_first64s(int n, int nSlices, int ld,
                                                        const int64_t* __restrict__ sliceOff,
                                                        const int* __restrict__ sliceW, int uw,
                                                        const int* __restrict__ perm, const uint16_t* __restrict__ loc,
                                                        const double* __restrict__ val, int stride,
                                                        const int* __restrict__ tileBase,
                                                        const uint16_t* __restrict__ tileCol,
                                                        const double* __restrict__ X, Row64* __restrict__ cA,
                                                        Row64* __restrict__ xa0, Row64* __restrict__ xa1,
                                                        double* __restrict__ out, FirstTermArgs ft) {
  extern __shared__ double2 ldsFirst[];
  const StagedRows<double> xs(ldsFirst, SP * kBlock);
  const int tile = blockIdx.x;
  const int slice = tile * kTileSlices + (int)threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const bool live = slice < nSlices;
  constexpr int S = kWave / G;
  const int sl = live ? slice : nSlices - 1;
  const int slot = sl * S + lane / G;
  const int row = perm ? perm[slot] : slot;
  const int sub = lane % G;
  const int width = UW > 0 ? UW : (uw > 0 ? uw : sliceW[sl]);
  const uint32_t off = (uint32_t)(UW > 0 || uw > 0 ? (int64_t)sl * width * kWave : sliceOff[sl]) +
                       (uint32_t)(UW == kSellLaneMajorW ? lane * width : lane);
  constexpr int CH = UW == kSellLaneMajorW ? kSellLaneMajorW : kChunk64;
  double v[CH];
  int li[CH];
  load_chunk<UW>(val, loc, off, 0, width, v, li);
  stage_tile_soa<SP>(xs, X, ld, tile, stride, tileBase, tileCol);
  __syncthreads();
  if (!live) return;
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
  for (int c0 = 0;;) {
#pragma unroll
    for (int u = 0; u < CH; u++) {
      const double2 a = xs.xy[li[u]];
      const double z = xs.z[li[u]];
      acc[0] += v[u] * a.x;
      acc[1] += v[u] * a.y;
      acc[2] += v[u] * z;
    }
    c0 += CH;
    if (c0 >= width) break;
    load_chunk<UW>(val, loc, off, c0, width, v, li);
  }
  group_reduce<G>(acc);
  double cr[3] = {0.0, 0.0, 0.0};
  if (row < n)
    for (int q = ft.off[row] + sub; q < ft.off[row + 1]; q += G) {
      const int j = ft.col[q];
      const double f = ft.val[q] / ft.D[j];
#pragma unroll
      for (int k = 0; k < 3; k++) cr[k] += f * ft.b[k * ld + j];
    }
  group_reduce<G>(cr);
  if (sub != 0 || row >= n) return;
  const double di = ft.D[row];
  double c[3], y[3], x0[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int idx = k * ld + row;
    c[k] = (ft.w * ft.w * cr[k] + (2.0 - ft.w) * ft.w * ft.b[idx]) / di;
    y[k] = acc[k] + c[k];
    x0[k] = X[idx];
  }
  cA[row] = Row64{c[0], c[1], c[2], 0.0};
  xa0[row] = Row64{x0[0], x0[1], x0[2], 0.0};
  xa1[row] = Row64{y[0], y[1], y[2], 0.0};
  if (out) {
    out[row] = y[0];
    out[ld + row] = y[1];
    out[2 * ld + row] = y[2];
  }
}

// rows of the staged union per lane: 2 or 4 (the union holds at most kStageMaxCols = 4 * kBlock)
static_assert(kStageMaxCols <= 4 * kBlock, "staged unions are loaded at most four rows per lane");
inline int stage_per_lane(int stride) { return stride <= 2 * kBlock ? 2 : 4; }
inline int tiles_of(const DeviceMesh& m) { return (m.nSlices + kTileSlices - 1) / kTileSlices; }

template <int G, int LAST>
void launch_sweep32s(const DeviceMesh& m, float W, const float4* cur, const float4* pre, float4* next, hipStream_t s) {
  const int tiles = tiles_of(m);
  const int sp = stage_per_lane(m.stageStride);
#define SW32(SP, UW)                                                                                               \
  launch_t(m.timer, TC_SWEEP32, k_sweep32s<G, LAST, SP, UW>, dim3(tiles), dim3(kBlock),                  \
           (uint32_t)staged_lds_bytes<float>(SP), s, m.nSim, m.nSlices, m.ld, W, m.sliceOff, m.sliceW,       \
           m.sellUniformW, m.sellPerm, m.m2Loc, m.m2ValF, m.stageStride, m.tileBase, m.tileCol, cur, pre, m.r0, next, \
           m.X)
  if (m.sellUniformW == kSellLaneMajorW) {  // lane-major slices
    if (sp == 2) SW32(2, kSellLaneMajorW);
    else SW32(4, kSellLaneMajorW);
  } else {
    if (sp == 2) SW32(2, 0);
    else SW32(4, 0);
  }
#undef SW32
}

template <int G>
void launch_sweep64s(const DeviceMesh& m, double W, const Row64* cur, const Row64* pre, Row64* next, double* out,
                     hipStream_t s) {
  const int tiles = tiles_of(m);
  const int sp = stage_per_lane(m.stageStride);
#define SW64(SP, UW)                                                                                               \
  launch_t(m.timer, TC_SWEEP64, k_sweep64s<G, SP, UW>, dim3(tiles), dim3(kBlock),                        \
           (uint32_t)staged_lds_bytes<double>(SP), s, m.nSim, m.nSlices, m.ld, W, m.sliceOff, m.sliceW,      \
           m.sellUniformW, m.sellPerm, m.m2Loc, m.m2Val, m.stageStride, m.tileBase, m.tileCol, cur, pre, m.cA, next, \
           out)
  if (m.sellUniformW == kSellLaneMajorW) {  // lane-major slices
    if (sp == 2) SW64(2, kSellLaneMajorW);
    else SW64(4, kSellLaneMajorW);
  } else {
    if (sp == 2) SW64(2, 0);
    else SW64(4, 0);
  }
#undef SW64
}

__global__ void k_apply_delta(int n, int ld, const float4* __restrict__ d, double* __restrict__ X) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 v = d[i];
  X[i] += (double)v.x;
  X[ld + i] += (double)v.y;
  X[2 * ld + i] += (double)v.z;
}

// The FP32 sweep with narrow staging: lane-major slices of width 8 at 8
// lanes per row, one workgroup per tile of kNarrowTileSlices slices, whose
// union (at most kNarrowMaxCols columns) gives every entry a one-byte index,
// so a lane's 8 entries are two 16-byte value loads and one 8-byte index
// load (6 -> 5 bytes per slot).  Otherwise k_sweep32s: the same sums in the
// same order, so bit-identical to it.
constexpr int kNarrowBlock = kNarrowTileSlices * kWave;
static_assert(kNarrowMaxCols <= 2 * kNarrowBlock, "narrow unions are staged at most two rows per lane");
template <int LAST>
__global__ void __launch_bounds__(kNarrowBlock) k_sweep32n(int n, int nSlices, int ld, float W,
                                                           const uint8_t* __restrict__ loc8,
                                                           const float* __restrict__ val, int stride,
                                                           const int* __restrict__ tileBase,
                                                           const uint16_t* __restrict__ tileCol,
                                                           const float4* __restrict__ dcur,
                                                           const float4* __restrict__ dpre,
                                                           const float4* __restrict__ r0, float4* __restrict__ dnext,
                                                           double* __restrict__ X) {
  constexpr int G = 8, S = kWave / G, UW = kSellLaneMajorW;
  __shared__ float2 sxy[kNarrowMaxCols];
  __shared__ float sz[kNarrowMaxCols];
  const int tile = blockIdx.x;
  const int slice = tile * kNarrowTileSlices + (int)threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const bool live = slice < nSlices;  // every lane stages and meets the barrier
  const int sl = live ? slice : nSlices - 1;
  const int row = sl * S + lane / G;  // uniform layouts keep the row order
  const uint32_t off = (uint32_t)sl * (uint32_t)(UW * kWave) + (uint32_t)(lane * UW);
  const float4 va = ld32((const float4*)val, off / 4), vb = ld32((const float4*)val, off / 4 + 1);
  const uint2 lw = ld32((const uint2*)loc8, off / 8);
  const float4 p = ld32(dpre, (uint32_t)row), r = ld32(r0, (uint32_t)row);
  {
    const int base = tileBase[tile];
    const uint32_t t0 = (uint32_t)tile * (uint32_t)stride;
    int c[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int u = threadIdx.x + q * kNarrowBlock;
      c[q] = base + (int)ld32(tileCol, t0 + (uint32_t)(u < stride ? u : stride - 1));
    }
    float4 g[2];
#pragma unroll
    for (int q = 0; q < 2; q++) g[q] = ld32(dcur, (uint32_t)c[q]);
#pragma unroll
    for (int q = 0; q < 2; q++) {
      sxy[threadIdx.x + q * kNarrowBlock] = make_float2(g[q].x, g[q].y);
      sz[threadIdx.x + q * kNarrowBlock] = g[q].z;
    }
  }
  __syncthreads();
  if (!live) return;
  const float v[8] = {va.x, va.y, va.z, va.w, vb.x, vb.y, vb.z, vb.w};
  const int li[8] = {(int)(lw.x & 0xff), (int)((lw.x >> 8) & 0xff), (int)((lw.x >> 16) & 0xff), (int)(lw.x >> 24),
                     (int)(lw.y & 0xff), (int)((lw.y >> 8) & 0xff), (int)((lw.y >> 16) & 0xff), (int)(lw.y >> 24)};
  float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int u = 0; u < 8; u++) {
    const float2 a = sxy[li[u]];
    const float z = sz[li[u]];
    acc[0] += v[u] * a.x;
    acc[1] += v[u] * a.y;
    acc[2] += v[u] * z;
  }
  group_reduce<G>(acc);
  if ((lane % G) != 0 || row >= n) return;
  float d0 = W * (acc[0] + r.x - p.x) + p.x;
  float d1 = W * (acc[1] + r.y - p.y) + p.y;
  float d2 = W * (acc[2] + r.z - p.z) + p.z;
  if (LAST) {
    X[row] += (double)d0;
    X[ld + row] += (double)d1;
    X[2 * ld + row] += (double)d2;
  } else {
    dnext[row] = make_float4(d0, d1, d2, 0.0f);
  }
}

// The FP64 sweep with narrow staging (as k_sweep32n: a lane's 8 entries are
// four 16-byte value loads and one 8-byte index load, 9 instead of 10 bytes
// per slot); otherwise k_sweep64s at width 8, so bit-identical to it.
__global__ void __launch_bounds__(kNarrowBlock) k_sweep64n(int n, int nSlices, int ld, double W,
                                                           const uint8_t* __restrict__ loc8,
                                                           const double* __restrict__ val, int stride,
                                                           const int* __restrict__ tileBase,
                                                           const uint16_t* __restrict__ tileCol,
                                                           const Row64* __restrict__ cur,
                                                           const Row64* __restrict__ pre,
                                                           const Row64* __restrict__ cA, Row64* __restrict__ next,
                                                           double* __restrict__ out) {
  constexpr int G = 8, S = kWave / G, UW = kSellLaneMajorW;
  __shared__ double2 sxy[kNarrowMaxCols];
  __shared__ double sz[kNarrowMaxCols];
  const int tile = blockIdx.x;
  const int slice = tile * kNarrowTileSlices + (int)threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const bool live = slice < nSlices;  // every lane stages and meets the barrier
  const int sl = live ? slice : nSlices - 1;
  const int row = sl * S + lane / G;  // uniform layouts keep the row order
  const uint32_t off = (uint32_t)sl * (uint32_t)(UW * kWave) + (uint32_t)(lane * UW);
  double2 va[4];
#pragma unroll
  for (int k = 0; k < 4; k++) va[k] = ld32((const double2*)val, off / 2 + k);
  const uint2 lw = ld32((const uint2*)loc8, off / 8);
  {
    const int base = tileBase[tile];
    const uint32_t t0 = (uint32_t)tile * (uint32_t)stride;
    int c[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int u = threadIdx.x + q * kNarrowBlock;
      c[q] = base + (int)ld32(tileCol, t0 + (uint32_t)(u < stride ? u : stride - 1));
    }
    Row64 g[2];
#pragma unroll
    for (int q = 0; q < 2; q++) g[q] = ld32(cur, (uint32_t)c[q]);
#pragma unroll
    for (int q = 0; q < 2; q++) {
      sxy[threadIdx.x + q * kNarrowBlock] = make_double2(g[q].x, g[q].y);
      sz[threadIdx.x + q * kNarrowBlock] = g[q].z;
    }
  }
  __syncthreads();
  if (!live) return;
  const double v[8] = {va[0].x, va[0].y, va[1].x, va[1].y, va[2].x, va[2].y, va[3].x, va[3].y};
  const int li[8] = {(int)(lw.x & 0xff), (int)((lw.x >> 8) & 0xff), (int)((lw.x >> 16) & 0xff), (int)(lw.x >> 24),
                     (int)(lw.y & 0xff), (int)((lw.y >> 8) & 0xff), (int)((lw.y >> 16) & 0xff), (int)(lw.y >> 24)};
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int u = 0; u < 8; u++) {
    const double2 a = sxy[li[u]];
    const double z = sz[li[u]];
    acc[0] += v[u] * a.x;
    acc[1] += v[u] * a.y;
    acc[2] += v[u] * z;
  }
  group_reduce<G>(acc);
  if ((lane % G) != 0 || row >= n) return;
  const Row64 p = ld32(pre, (uint32_t)row), c = ld32(cA, (uint32_t)row);
  double x0 = W * (acc[0] + c.x - p.x) + p.x;
  double x1 = W * (acc[1] + c.y - p.y) + p.y;
  double x2 = W * (acc[2] + c.z - p.z) + p.z;
  next[row] = Row64{x0, x1, x2, 0.0};
  if (out) {
    out[row] = x0;
    out[ld + row] = x1;
    out[2 * ld + row] = x2;
  }
}

template <int G, int LAST>
void launch_sweep32(const DeviceMesh& m, int grid, float W, const float4* cur, const float4* pre, float4* next,
                    hipStream_t s) {
  if constexpr (G == 8) {
    if (m.m2Loc8) {
      launch_t(m.timer, TC_SWEEP32, k_sweep32n<LAST>, dim3((m.nSlices + kNarrowTileSlices - 1) / kNarrowTileSlices),
               dim3(kNarrowBlock), 0, s, m.nSim, m.nSlices, m.ld, W, m.m2Loc8, m.m2ValF, m.narrowStride,
               m.ntileBase, m.ntileCol, cur, pre, m.r0, next, m.X);
      return;
    }
  }
  if constexpr (G > 1) {  // one lane per row is never staged (its tiles span 256 rows)
    if (m.m2Loc) {
      launch_sweep32s<G, LAST>(m, W, cur, pre, next, s);
      return;
    }
  }
  launch_t(m.timer, TC_SWEEP32, m.m2Col16 ? k_sweep32<G, LAST, true> : k_sweep32<G, LAST, false>, dim3(grid),
           dim3(kBlock), 0, s, m.nSim, m.nSlices, m.ld, W, m.sliceOff, m.sliceW, m.sellUniformW, m.sellPerm, m.m2Col,
           m.m2Col16, m.m2ValF, cur, pre, m.r0, next, m.X);
}

// The FP64 solve: x_0 = X, iterates rotate through xa0..xa2, the last one is
// also written to X (SoA).  outs (when given: the recorded solve of the adjoint's
// forward pass) receives every iterate x_1 .. x_iters in SoA as well; both
// forms run the same kernels, so their iterates are bit-identical.
template <int G>
void solve64(const DeviceMesh& m, double w, double rho, int iters, double* const* outs, hipStream_t s) {
  const int grid = (m.nSlices * kWave + kBlock - 1) / kBlock;
  const FirstTermArgs ft{m.r1Off, m.r1Col, m.r1Val, m.sysDiag, m.sysRhs, w};
  double r = rho * rho;
  double W = 2.0 / (2.0 - r);
  const bool c16 = m.m2Col16 != nullptr;
  bool stagedFirst = false;
  if constexpr (G > 1) {  // one lane per row is never staged
    if (m.m2Loc) {
      const int sp = stage_per_lane(m.stageStride);
#define F64(SP, UW)                                                                                                \
  launch_t(m.timer, TC_FIRST, k_first64s<G, SP, UW>, dim3(tiles_of(m)), dim3(kBlock),                              \
           (uint32_t)staged_lds_bytes<double>(SP), s, m.nSim, m.nSlices, m.ld, m.sliceOff, m.sliceW, m.sellUniformW, \
           m.sellPerm, m.m2Loc, m.m2Val, m.stageStride, m.tileBase, m.tileCol, m.X, m.cA, m.xa0, m.xa1,              \
           outs ? outs[1] : nullptr, ft)
      if (m.sellUniformW == kSellLaneMajorW) {  // lane-major slices
        if (sp == 2) F64(2, kSellLaneMajorW);
        else F64(4, kSellLaneMajorW);
      } else {
        if (sp == 2) F64(2, 0);
        else F64(4, 0);
      }
#undef F64
      stagedFirst = true;
    }
  }
  if (!stagedFirst)
    launch_t(m.timer, TC_FIRST, c16 ? k_first64<G, true> : k_first64<G, false>, dim3(grid), dim3(kBlock), 0, s,
             m.nSim, m.nSlices, m.ld, m.sliceOff, m.sliceW, m.sellUniformW, m.sellPerm, m.m2Col, m.m2Col16, m.m2Val,
             m.X, m.cA, m.xa0, m.xa1, outs ? outs[1] : nullptr, ft);
  Row64 *pre = m.xa0, *cur = m.xa1, *next = m.xa2;
  if (iters <= 1) {
    hipLaunchKernelGGL(k_row64_to_soa, dim3((m.nSim + kBlock - 1) / kBlock), dim3(kBlock), 0, s, m.nSim, m.ld, cur,
                       m.X);
    return;
  }
  for (int it = 1; it < iters; it++) {
    double* out = outs ? outs[it + 1] : (it == iters - 1 ? m.X : nullptr);
    bool staged = false;
    if constexpr (G == 8) {
      if (m.m2Loc8) {
        launch_t(m.timer, TC_SWEEP64, k_sweep64n,
                 dim3((m.nSlices + kNarrowTileSlices - 1) / kNarrowTileSlices), dim3(kNarrowBlock), 0, s, m.nSim,
                 m.nSlices, m.ld, W, m.m2Loc8, m.m2Val, m.narrowStride, m.ntileBase, m.ntileCol, cur, pre, m.cA,
                 next, out);
        staged = true;
      }
    }
    if constexpr (G > 1) {
      if (!staged && m.m2Loc) {
        launch_sweep64s<G>(m, W, cur, pre, next, out, s);
        staged = true;
      }
    }
    if (!staged)
      launch_t(m.timer, TC_SWEEP64, c16 ? k_sweep64<G, true> : k_sweep64<G, false>, dim3(grid), dim3(kBlock), 0, s,
               m.nSim, m.nSlices, m.ld, W, m.sliceOff, m.sliceW, m.sellUniformW, m.sellPerm, m.m2Col, m.m2Col16,
               m.m2Val, cur, pre, m.cA, next, out);
    W = 4.0 / (4.0 - r * W);
    Row64* t = pre; pre = cur; cur = next; next = t;
  }
  // the sweeps write the simulated rows only: copy those back (the static rows of X stay)
  if (outs) launch_axpby(m, 1.0, outs[iters], 0.0, nullptr, m.X, s);
}

template <int G>
void solve_impl(const DeviceMesh& m, double w, double rho, int iters, int precision, bool residualReady,
                hipStream_t s) {
  if (precision == 0) {
    solve64<G>(m, w, rho, iters, nullptr, s);
    return;
  }
  const int grid = (m.nSlices * kWave + kBlock - 1) / kBlock;
  double r = rho * rho;
  double W = 2.0 / (2.0 - r);
  float4 *pre = m.dPre, *cur = m.dCur, *next = m.dNext;
  // delta_0 = 0 -> pre, delta_1 = r0 -> cur
  constexpr int GR = 4;  // lanes per R1 row
  const int gridR = (m.nSim * GR + kBlock - 1) / kBlock;
  if (!residualReady)  // else the local step's k_rhs wrote v and pt
    launch_t(m.timer, TC_FIRST, k_res_a<GR>, dim3(gridR), dim3(kBlock), 0, s, m.nSim, m.ld, w, m.r1Off, m.r1Col,
             m.r1Val, m.sysDiag, m.X, m.sysRhs, m.xa0, m.cA);
  launch_t(m.timer, TC_FIRST, k_res_b<GR>, dim3(gridR), dim3(kBlock), 0, s, m.nSim, w, m.r1Off, m.r1Col, m.r1Val,
           m.sysDiag, m.xa0, m.cA, pre, cur, m.r0);
  if (iters <= 1) {
    hipLaunchKernelGGL(k_apply_delta, dim3((m.nSim + kBlock - 1) / kBlock), dim3(kBlock), 0, s, m.nSim, m.ld, cur,
                       m.X);
    return;
  }
  for (int it = 1; it < iters; it++) {
    if (it == iters - 1)
      launch_sweep32<G, 1>(m, grid, (float)W, cur, pre, next, s);
    else
      launch_sweep32<G, 0>(m, grid, (float)W, cur, pre, next, s);
    W = 4.0 / (4.0 - r * W);
    float4* t = pre; pre = cur; cur = next; next = t;
  }
}

// Reverse sweeps of the Chebyshev recurrence in the scaled adjoint u = D^{-1} xbar
// (M^T = D M D^{-1} because R1 is symmetric, so M^T xbar = D M u): each is the
// forward SELL SpMV on u,
//   u_out = a M u_in + b u_oth,   uc[row] += a u_in[row]   (uc: D^{-1}-scaled first-term adjoint);
// the M^2-entry adjoints are gathered once per solve by k_adj_gbar.
template <int G, bool C16>
__global__ void __launch_bounds__(kBlock, 8) k_adj_spmv(int n, int nSlices, int ld, double a, double b,
                                                        const int64_t* __restrict__ sliceOff,
                                                        const int* __restrict__ sliceW, int uw, const int* __restrict__ perm,
                                                        const int* __restrict__ col, const int16_t* __restrict__ col16,
                                                        const double* __restrict__ val, const double* __restrict__ uin,
                                                        const double* __restrict__ uoth, double* __restrict__ uout,
                                                        double* __restrict__ uc) {
  int gt = blockIdx.x * blockDim.x + threadIdx.x;
  int slice = gt / kWave, lane = gt % kWave;
  if (slice >= nSlices) return;
  constexpr int S = kWave / G;
  const int slot = slice * S + lane / G;
  const int row = perm ? perm[slot] : slot;
  int sub = lane % G;
  const int width = uw > 0 ? uw : sliceW[slice];
  const bool lm = uw == kSellLaneMajorW;  // lane-major slices, else chunk-major
  const int64_t off = (uw > 0 ? (int64_t)slice * uw * kWave : sliceOff[slice]) + (lm ? (int64_t)lane * width : lane);
  const int step = lm ? 1 : kWave;
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll 4
  for (int c = 0; c < width; c++) {
    int64_t e = off + (int64_t)c * step;
    int j = C16 ? slot + (int)col16[e] : col[e];
    double v = val[e];
    acc[0] += v * uin[j];
    acc[1] += v * uin[ld + j];
    acc[2] += v * uin[2 * ld + j];
  }
  group_reduce<G>(acc);
  if (sub != 0 || row >= n) return;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int idx = k * ld + row;
    uout[idx] = a * acc[k] + (b != 0.0 ? b * uoth[idx] : 0.0);
    uc[idx] += a * uin[idx];
  }
}

// Gbar_e += sum_k W_k u_{k+1}[row] . x_k[col] over the K sweeps of one pass
// (u_{k+1} and x_k from the pass's reverse and recorded forward iterates;
// x_0, and x_k = x1 + (k-1) xStride for k >= 1).  One workgroup per slice, one
// lane per column of the union of the slice's columns (sorted, listed per slice
// at set-up): the lane loads its column's x_k kGbarChunk sweeps at a time (all
// loads of a chunk in flight together) and sums over k, in order, for all R
// rows of the slice at once; the rows' u values are wave-uniform (scalar
// loads: R consecutive doubles per component, or with P the rows that the
// slice's slots hold under the SELL row order perm; rows past n read the
// vectors' ld padding, and their sums are never stored).  So each x value is read once per
// slice, where a lane per entry reads it for every entry of the column.  uPos
// gives, per union column and row, the entry's offset in the slice's SELL block
// (0xFFFF: no entry; padding slots are never written).
struct GbarArgs {
  const double* u[kMaxCheb];  // u_1 .. u_K
  double w[kMaxCheb];
  int K;
};
constexpr int kGbarThreads = 128;
constexpr int kGbarChunk = 8;
template <int G, bool P>
__global__ void __launch_bounds__(kGbarThreads) k_adj_gbar(int ld, const int64_t* __restrict__ sliceOff,
                                                           const int* __restrict__ perm,
                                                           const int* __restrict__ uOff, const int* __restrict__ uCol,
                                                           const uint16_t* __restrict__ uPos,
                                                           const double* __restrict__ x0, const double* __restrict__ x1,
                                                           int64_t xStride, GbarArgs ga, double* __restrict__ Gbar) {
  constexpr int R = kWave / G;
  const int slice = blockIdx.x;
  const int K = ga.K;
  const int row0 = slice * R;
  int rows[R];  // the slice's rows, relative to row0 without P
#pragma unroll
  for (int i = 0; i < R; i++) rows[i] = P ? perm[row0 + i] : i;
  const int u0 = uOff[slice], nU = uOff[slice + 1] - u0;
  const int64_t sb = sliceOff[slice];
  for (int t = threadIdx.x; t < nU; t += kGbarThreads) {
    const int j = uCol[u0 + t];
    double g[R];
#pragma unroll
    for (int i = 0; i < R; i++) g[i] = 0.0;
    for (int k0 = 0; k0 < K; k0 += kGbarChunk) {
      double xv[kGbarChunk][3];
#pragma unroll
      for (int q = 0; q < kGbarChunk; q++)
        if (k0 + q < K) {
          const int k = k0 + q;
          const double* x = k == 0 ? x0 : x1 + (k - 1) * xStride;
          xv[q][0] = x[j];
          xv[q][1] = x[ld + j];
          xv[q][2] = x[2 * ld + j];
        }
#pragma unroll
      for (int q = 0; q < kGbarChunk; q++)
        if (k0 + q < K) {
          const int k = k0 + q;
          const double w = ga.w[k];
          const double* uk = ga.u[k] + (P ? 0 : row0);
#pragma unroll
          for (int i = 0; i < R; i++)
            g[i] += w * (uk[rows[i]] * xv[q][0] + uk[ld + rows[i]] * xv[q][1] + uk[2 * ld + rows[i]] * xv[q][2]);
        }
    }
    const uint16_t* pos = uPos + (int64_t)(u0 + t) * R;
#pragma unroll
    for (int i = 0; i < R; i++) {
      const int p = pos[i];
      if (p != 0xFFFF) Gbar[sb + p] += g[i];
    }
  }
}

// The reverse sweep with its operand staged in LDS (as k_sweep64s; the
// reverse iterates are SoA, so a staged row is three loads)
template <int G, int SP, int UW>
__global__ void __launch_bounds__(kBlock, 8) k_adj_spmv_s(int n, int nSlices, int ld, double a, double b,
                                                          const int64_t* __restrict__ sliceOff,
                                                          const int* __restrict__ sliceW, int uw,
                                                          const int* __restrict__ perm,
                                                          const uint16_t* __restrict__ loc,
                                                          const double* __restrict__ val, int stride,
                                                          const int* __restrict__ tileBase,
                                                          const uint16_t* __restrict__ tileCol,
                                                          const double* __restrict__ uin,
                                                          const double* __restrict__ uoth, double* __restrict__ uout,
                                                          double* __restrict__ uc) {
  extern __shared__ double2 ldsAdj[];
  const StagedRows<double> us(ldsAdj, SP * kBlock);
  const int tile = blockIdx.x;
  const int slice = tile * kTileSlices + (int)threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const bool live = slice < nSlices;
  constexpr int S = kWave / G;
  const int sl = live ? slice : nSlices - 1;
  const int slot = sl * S + lane / G;
  const int row = perm ? perm[slot] : slot;
  const int width = UW > 0 ? UW : (uw > 0 ? uw : sliceW[sl]);
  const uint32_t off = (uint32_t)(UW > 0 || uw > 0 ? (int64_t)sl * width * kWave : sliceOff[sl]) +
                       (uint32_t)(UW == kSellLaneMajorW ? lane * width : lane);
  constexpr int CH = UW == kSellLaneMajorW ? kSellLaneMajorW : kChunk64;
  double v[CH];
  int li[CH];
  load_chunk<UW>(val, loc, off, 0, width, v, li);
  stage_tile_soa<SP>(us, uin, ld, tile, stride, tileBase, tileCol);
  __syncthreads();
  if (!live) return;
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
  for (int c0 = 0;;) {
#pragma unroll
    for (int u = 0; u < CH; u++) {
      const double2 xy = us.xy[li[u]];
      const double z = us.z[li[u]];
      acc[0] += v[u] * xy.x;
      acc[1] += v[u] * xy.y;
      acc[2] += v[u] * z;
    }
    c0 += CH;
    if (c0 >= width) break;
    load_chunk<UW>(val, loc, off, c0, width, v, li);
  }
  group_reduce<G>(acc);
  if ((lane % G) != 0 || row >= n) return;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int idx = k * ld + row;
    uout[idx] = a * acc[k] + (b != 0.0 ? b * uoth[idx] : 0.0);
    uc[idx] += a * uin[idx];
  }
}

template <int G>
void adj_pass_impl(const DeviceMesh& m, const AdjMesh& a, const double* W, int K, double* const* u, const double* x0,
                   const double* x1, int64_t xStride, double* Gbar, double* uc, hipStream_t s) {
  const int grid = (m.nSlices * kWave + kBlock - 1) / kBlock;
  for (int k = K - 1; k >= 0; k--) {
    double b = k + 1 <= K - 1 ? 1.0 - W[k + 1] : 0.0;
    bool staged = false;
    if constexpr (G > 1) {
      if (m.m2Loc) {
        const int tiles = tiles_of(m);
        const int sp = stage_per_lane(m.stageStride);
#define ADJ(SP, UW)                                                                                             \
  hipLaunchKernelGGL((k_adj_spmv_s<G, SP, UW>), dim3(tiles), dim3(kBlock),                            \
                     staged_lds_bytes<double>(SP), s, m.nSim, m.nSlices, m.ld, W[k], b, m.sliceOff,     \
                     m.sliceW, m.sellUniformW, m.sellPerm, m.m2Loc, m.m2Val, m.stageStride, m.tileBase,         \
                     m.tileCol, u[k + 1], u[k + 2], u[k], uc)
        if (m.sellUniformW == kSellLaneMajorW) {  // lane-major slices
          if (sp == 2) ADJ(2, kSellLaneMajorW);
          else ADJ(4, kSellLaneMajorW);
        } else {
          if (sp == 2) ADJ(2, 0);
          else ADJ(4, 0);
        }
#undef ADJ
        staged = true;
      }
    }
    if (!staged)
      hipLaunchKernelGGL((m.m2Col16 ? k_adj_spmv<G, true> : k_adj_spmv<G, false>), dim3(grid), dim3(kBlock), 0, s,
                         m.nSim, m.nSlices, m.ld, W[k], b, m.sliceOff, m.sliceW, m.sellUniformW, m.sellPerm, m.m2Col,
                         m.m2Col16, m.m2Val, u[k + 1], u[k + 2], u[k], uc);
  }
  GbarArgs ga;
  ga.K = K;
  for (int k = 0; k < K; k++) {
    ga.u[k] = u[k + 1];
    ga.w[k] = W[k];
  }
  hipLaunchKernelGGL((m.sellPerm ? k_adj_gbar<G, true> : k_adj_gbar<G, false>), dim3(m.nSlices), dim3(kGbarThreads),
                     0, s, m.ld, m.sliceOff, m.sellPerm, a.uOff, a.uCol, a.uPos, x0, x1, xStride, ga, Gbar);
}

// Adjoint of the first term c_i = (w^2 sum_j R1_ij b_j / D_j + (2 - w) w b_i) / D_i
// given its scaled adjoint uc = D^{-1} cbar:  bbar = w^2 D^{-1} R1 uc + (2 - w) w uc,
// and the D and R1 parts of the derivative accumulated into Dbar and R1bar.
// kFirstLanes lanes per row share its R1 entries; rows past n run on row n - 1
// without writing, so every lane takes part in the reduction.
constexpr int kFirstLanes = 8;
__global__ void __launch_bounds__(kBlock) k_adj_first(int n, int ld, double w, const int* __restrict__ off,
                                                      const int* __restrict__ col, const double* __restrict__ r1,
                                                      const double* __restrict__ D, const double* __restrict__ b,
                                                      const double* __restrict__ uc, double* __restrict__ bbar,
                                                      double* __restrict__ Dbar, double* __restrict__ R1bar) {
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int sub = gt % kFirstLanes;
  const bool live = gt / kFirstLanes < n;
  const int i = live ? gt / kFirstLanes : n - 1;
  double ui[3] = {uc[i], uc[ld + i], uc[2 * ld + i]};
  double s1[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
  const double w2 = w * w;
  for (int q = off[i] + sub; q < off[i + 1]; q += kFirstLanes) {
    int j = col[q];
    double a = r1[q], dj = D[j];
    double bj[3] = {b[j] / dj, b[ld + j] / dj, b[2 * ld + j] / dj};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      s1[k] += a * uc[k * ld + j];
      s2[k] += a * bj[k];
    }
    if (live) R1bar[q] += w2 * (ui[0] * bj[0] + ui[1] * bj[1] + ui[2] * bj[2]);
  }
  group_reduce<kFirstLanes>(s1);
  group_reduce<kFirstLanes>(s2);
  if (!live || sub != 0) return;
  double di = D[i];
  double db = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double bi = b[k * ld + i];
    double c = (w2 * s2[k] + (2.0 - w) * w * bi) / di;
    bbar[k * ld + i] = w2 * s1[k] / di + (2.0 - w) * w * ui[k];
    db -= ui[k] * c + w2 * bi * s1[k] / (di * di);
  }
  Dbar[i] += db;
}

// Adjoint of the operator update as three gathers (no atomics, fixed order):
//   rows: Dbar_i -= sum_e Gbar_e (w^2 s_e + 2 w (1-w) R1_e) / D_i over row i's
//         entries (s_e = sum_paths A_p / D_k), and R1bar at each entry's own
//         R1 position += 2 w (1-w) Gbar_e;
//   mid:  Dbar_k -= (w^2 / D_k^2) sum_{paths through k} Gbar_e A_p;
//   r1:   R1bar_q += sum_{paths using q} (w^2 / D_k) Gbar_e R1_other (spring entries only,
//         the R1 values the stiffness gradient reads).
// wave-wide sum in a fixed order (xor tree): deterministic
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// one wavefront per row; lane t takes the row's entries t, t + 64, ...
// (slotOf: the row's SELL slot when rows are reordered)
template <int G>
__global__ void __launch_bounds__(kBlock) k_adj_ops_rows(int n, const int64_t* __restrict__ sliceOff,
                                                         const int* __restrict__ sliceW, int uw,
                                                         const int* __restrict__ slotOf, const double* __restrict__ r1e,
                                                         const int* __restrict__ r1Idx,
                                                         const int64_t* __restrict__ pOff, const int* __restrict__ pK,
                                                         const double* __restrict__ pA, const double* __restrict__ D,
                                                         const double* __restrict__ Gbar, double w,
                                                         double* __restrict__ Dbar, double* __restrict__ R1bar) {
  int row = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  int lane = threadIdx.x % kWave;
  if (row >= n) return;  // whole wavefront
  constexpr int S = kWave / G;
  const int slot = slotOf ? slotOf[row] : row;
  int slice = slot / S, lr = slot % S;
  const int64_t base = sliceOff[slice];
  const int width = sliceW[slice];
  int cnt = width * G;
  const double w2 = w * w, c1 = 2.0 * (1.0 - w) * w;
  double acc = 0.0;
  const bool lm = uw == kSellLaneMajorW;  // lane-major slices, else chunk-major
  for (int t = lane; t < cnt; t += kWave) {
    // entry t of the row: lane lr*G + t%G, chunk t/G
    const int64_t e = base + (lm ? (int64_t)(lr * G + t % G) * width + t / G : (int64_t)(t / G) * kWave + lr * G + t % G);
    double g = Gbar[e];
    if (g == 0.0) continue;
    double sp = 0.0;
    for (int64_t p = pOff[e]; p < pOff[e + 1]; p++) sp += pA[p] / D[pK[p]];
    acc += g * (w2 * sp + c1 * r1e[e]);
    int q = r1Idx[e];
    if (q >= 0) R1bar[q] += c1 * g;
  }
  acc = wave_sum(acc);
  if (lane == 0) Dbar[row] -= acc / D[row];
}

// one wavefront per middle point k
__global__ void __launch_bounds__(kBlock) k_adj_ops_mid(int n, const int* __restrict__ off,
                                                        const int64_t* __restrict__ ent, const double* __restrict__ a,
                                                        const double* __restrict__ D, const double* __restrict__ Gbar,
                                                        double w, double* __restrict__ Dbar) {
  int k = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  int lane = threadIdx.x % kWave;
  if (k >= n) return;
  double acc = 0.0;
  for (int q = off[k] + lane; q < off[k + 1]; q += kWave) acc += Gbar[ent[q]] * a[q];
  acc = wave_sum(acc);
  double dk = D[k];
  if (lane == 0) Dbar[k] -= w * w * acc / (dk * dk);
}

// one wavefront per spring R1 entry
__global__ void __launch_bounds__(kBlock) k_adj_ops_r1(int nq, const int* __restrict__ qIdx,
                                                       const int* __restrict__ off, const int64_t* __restrict__ ent,
                                                       const int* __restrict__ mid, const int* __restrict__ other,
                                                       const double* __restrict__ r1Val, const double* __restrict__ D,
                                                       const double* __restrict__ Gbar, double w,
                                                       double* __restrict__ R1bar) {
  int t = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  int lane = threadIdx.x % kWave;
  if (t >= nq) return;
  double acc = 0.0;
  for (int q = off[t] + lane; q < off[t + 1]; q += kWave) acc += Gbar[ent[q]] * r1Val[other[q]] / D[mid[q]];
  acc = wave_sum(acc);
  if (lane == 0) R1bar[qIdx[t]] += w * w * acc;
}

template <int G>
void one_sweep_impl(const DeviceMesh& m, int precision, hipStream_t s) {
  const int grid = (m.nSlices * kWave + kBlock - 1) / kBlock;
  const double W = 1.9;
  if constexpr (G > 1) {
    if (precision == 0 && m.m2Loc) {
      launch_sweep64s<G>(m, W, m.xa1, m.xa0, m.xa2, nullptr, s);
      return;
    }
  }
  if (precision == 0)
    hipLaunchKernelGGL((m.m2Col16 ? k_sweep64<G, true> : k_sweep64<G, false>), dim3(grid), dim3(kBlock), 0, s,
                       m.nSim, m.nSlices, m.ld, W, m.sliceOff, m.sliceW, m.sellUniformW, m.sellPerm, m.m2Col, m.m2Col16, m.m2Val,
                       m.xa1, m.xa0, m.cA, m.xa2, nullptr);
  else
    launch_sweep32<G, 0>(m, grid, (float)W, m.dCur, m.dPre, m.dNext, s);
}

}  // namespace

void launch_one_sweep(const DeviceMesh& m, int precision, hipStream_t s) {
  switch (m.sellLanes) {
    case 1: one_sweep_impl<1>(m, precision, s); break;
    case 2: one_sweep_impl<2>(m, precision, s); break;
    case 8: one_sweep_impl<8>(m, precision, s); break;
    default: one_sweep_impl<4>(m, precision, s); break;
  }
}

void launch_operators(const DeviceMesh& m, double omega, int precision, hipStream_t s) {
  (void)precision;  // both the FP64 and the FP32 copies are kept current
  int64_t nE = m.sellEntries;
  launch_t(m.timer, TC_OPERATORS, k_dirty_rows, dim3((m.nSim + kBlock - 1) / kBlock), dim3(kBlock), 0, s, m.nSim,
           m.r1Off, m.r1Col, m.dChanged, m.dDirty);
  launch_t(m.timer, TC_OPERATORS, k_operators, dim3((unsigned)((nE + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, nE,
           omega, m.m2RowFlag, m.dDirty, m.m2R1, m.p2Off, m.p2K, m.p2A, m.sysDiag, m.m2Val, m.m2ValF);
}

void launch_solve_taped(const DeviceMesh& m, double omega, double rho, int iters, double* const* outs, hipStream_t s) {
  switch (m.sellLanes) {
    case 1: solve64<1>(m, omega, rho, iters, outs, s); break;
    case 2: solve64<2>(m, omega, rho, iters, outs, s); break;
    case 8: solve64<8>(m, omega, rho, iters, outs, s); break;
    default: solve64<4>(m, omega, rho, iters, outs, s); break;
  }
}

void launch_adj_pass(const DeviceMesh& m, const AdjMesh& a, const double* W, int K, double* const* u,
                     const double* x0, const double* x1, int64_t xStride, double* Gbar, double* uc, hipStream_t s) {
  switch (m.sellLanes) {
    case 1: adj_pass_impl<1>(m, a, W, K, u, x0, x1, xStride, Gbar, uc, s); break;
    case 2: adj_pass_impl<2>(m, a, W, K, u, x0, x1, xStride, Gbar, uc, s); break;
    case 8: adj_pass_impl<8>(m, a, W, K, u, x0, x1, xStride, Gbar, uc, s); break;
    default: adj_pass_impl<4>(m, a, W, K, u, x0, x1, xStride, Gbar, uc, s); break;
  }
}

void launch_adj_first(const DeviceMesh& m, double omega, const double* b, const double* uc, double* bbar, double* Dbar,
                      double* R1bar, hipStream_t s) {
  const int rowsPerBlock = kBlock / kFirstLanes;
  hipLaunchKernelGGL(k_adj_first, dim3((m.nSim + rowsPerBlock - 1) / rowsPerBlock), dim3(kBlock), 0, s, m.nSim, m.ld,
                     omega, m.r1Off, m.r1Col, m.r1Val, m.sysDiag, b, uc, bbar, Dbar, R1bar);
}

void launch_adj_ops_gather(const DeviceMesh& m, const AdjMesh& a, double omega, const double* Gbar, double* Dbar,
                           double* R1bar, hipStream_t s) {
  const int rowsPerBlock = kBlock / kWave;
  const int gr = (m.nSim + rowsPerBlock - 1) / rowsPerBlock;
#define ROWS(G)                                                                                                    \
  hipLaunchKernelGGL(k_adj_ops_rows<G>, dim3(gr), dim3(kBlock), 0, s, m.nSim, m.sliceOff, m.sliceW, m.sellUniformW, m.sellSlotOf, m.m2R1,        \
                     a.sellR1Idx, m.p2Off, m.p2K, m.p2A, m.sysDiag, Gbar, omega, Dbar, R1bar)
  switch (m.sellLanes) {
    case 1: ROWS(1); break;
    case 2: ROWS(2); break;
    case 8: ROWS(8); break;
    default: ROWS(4); break;
  }
#undef ROWS
  hipLaunchKernelGGL(k_adj_ops_mid, dim3(gr), dim3(kBlock), 0, s, m.nSim, a.midOff, a.midEnt, a.midA, m.sysDiag, Gbar,
                     omega, Dbar);
  if (a.nSpringQ > 0)
    hipLaunchKernelGGL(k_adj_ops_r1, dim3((a.nSpringQ + rowsPerBlock - 1) / rowsPerBlock), dim3(kBlock), 0, s,
                       a.nSpringQ, a.springQ, a.springOff, a.springEnt, a.springMid, a.springOther, m.r1Val,
                       m.sysDiag, Gbar, omega, R1bar);
}

void launch_solve(const DeviceMesh& m, double omega, double rho, int iters, int precision, hipStream_t s,
                  bool residualReady) {
  switch (m.sellLanes) {
    case 1: solve_impl<1>(m, omega, rho, iters, precision, residualReady, s); break;
    case 2: solve_impl<2>(m, omega, rho, iters, precision, residualReady, s); break;
    case 8: solve_impl<8>(m, omega, rho, iters, precision, residualReady, s); break;
    default: solve_impl<4>(m, omega, rho, iters, precision, residualReady, s); break;
  }
}

}  // namespace pdipc
