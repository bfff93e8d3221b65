// Broad phase (gpuCcdCulling, Simulator/PD-IPC/PdIpcSimulator.cpp:1799-1869), rebuilt
// MI355X-first:
//   1. swept, inflated patch boxes (updateCcdPatchBboxesKernel, PdIpcGpuFunc.cu:1407-1473);
//   2. a Karras LBVH over the patch boxes, rebuilt every step on the device:
//      30-bit Morton codes of the box centroids normalised to the scene box
//      (lbvh/morton_code.cuh:9-40, bvh.cuh:197-221), a stable LSD radix sort
//      (8-bit digits; in-block stable ranking with wavefront ballots), the
//      (code << 32 | index) tie-break, Karras determine_range / find_split
//      (bvh.cuh:69-163); node boxes from a sparse table of leaf-box unions
//      (no inter-block synchronisation), stored rounded outward to float with
//      preorder skip links;
//   3. every patch, in Morton order, finds each patch q > p whose box overlaps
//      (checkOverlap, PdIpcGpuFunc.cu:1615-1626, closed boxes, confirmed in
//      FP64) in a two-pass stackless traversal (top of the tree, then one lane
//      per overlapping small subtree); every wavefront appends to a segment
//      of its own and the segments are compacted in wavefront order, so the
//      pair list (and the record numbering after it) is reproducible;
//   4. every overlapping patch pair expands to its point-triangle and
//      edge-edge records (getCcdPotentialICPairsListKernel,
//      PdIpcGpuFunc.cu:1838-1948), one lane per record: all point-triangle
//      records first (pair order, the reference's order within a pair), then
//      all edge-edge records.
// The pair set equals the reference heap-BVH's (both are "all overlapping
// pairs"); the heap topology itself is not reproduced.
#include "device.hpp"

namespace pdipc {
namespace {

constexpr int kBlock = 256;
constexpr int kScanItems = 4;                       // per thread
constexpr int kScanTile = kBlock * kScanItems;      // 1024
constexpr int kRadixItems = 8;
constexpr int kRadixTile = kBlock * kRadixItems;    // 2048

__device__ inline double dmin(double a, double b) { return a < b ? a : b; }
__device__ inline double dmax(double a, double b) { return a > b ? a : b; }

// ---------------- patch boxes ----------------
// kBoxLanes lanes per patch, each taking every kBoxLanes-th face, then a
// min / max across the patch's lanes: exact and order-free, so the boxes are
// those of one lane walking all the faces (one lane per patch left the
// kernel at ~1.5 wavefronts per CU: 41 us for the 95 k patches of a 256^2 cloth over a
// subdivision-6 icosphere).
constexpr int kBoxLanes = 8;
__global__ void k_patch_boxes(int P, int ld, const int* __restrict__ pfOff, const int* __restrict__ pfList,
                              const int3* __restrict__ faces, const double* __restrict__ X,
                              const double* __restrict__ DX, int nSim, double infl, double* __restrict__ box) {
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = gt / kBoxLanes, sub = gt % kBoxLanes;
  double lo[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
  double hi[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
  if (p < P)
    for (int q = pfOff[p] + sub; q < pfOff[p + 1]; q += kBoxLanes) {
      int3 f = faces[pfList[q]];
      int v[3] = {f.x, f.y, f.z};
#pragma unroll
      for (int k = 0; k < 3; k++) {
        double a0 = X[k * ld + v[0]], a1 = X[k * ld + v[1]], a2 = X[k * ld + v[2]];
        double b0 = a0 + (v[0] < nSim ? DX[k * ld + v[0]] : 0.0);
        double b1 = a1 + (v[1] < nSim ? DX[k * ld + v[1]] : 0.0);
        double b2 = a2 + (v[2] < nSim ? DX[k * ld + v[2]] : 0.0);
        double mn = dmin(dmin(dmin(a0, a1), dmin(a2, b0)), dmin(b1, b2)) - infl;
        double mx = dmax(dmax(dmax(a0, a1), dmax(a2, b0)), dmax(b1, b2)) + infl;
        lo[k] = dmin(lo[k], mn);
        hi[k] = dmax(hi[k], mx);
      }
    }
#pragma unroll
  for (int o = kBoxLanes / 2; o > 0; o >>= 1)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      lo[k] = dmin(lo[k], __shfl_xor(lo[k], o, kWave));
      hi[k] = dmax(hi[k], __shfl_xor(hi[k], o, kWave));
    }
  if (p >= P || sub != 0) return;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    box[k * P + p] = lo[k];
    box[(3 + k) * P + p] = hi[k];
  }
}

// ---------------- scene box of the leaf boxes (lbvh's aabb_whole) ----------------
__global__ void k_scene_partial(int P, const double* __restrict__ box, double* __restrict__ part) {
  __shared__ double red[6][kBlock / kWave];
  double v[6];
#pragma unroll
  for (int k = 0; k < 3; k++) { v[k] = 1.7976931348623157e308; v[3 + k] = -1.7976931348623157e308; }
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      v[k] = dmin(v[k], box[k * P + p]);
      v[3 + k] = dmax(v[3 + k], box[(3 + k) * P + p]);
    }
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      v[k] = dmin(v[k], __shfl_xor(v[k], o, kWave));
      v[3 + k] = dmax(v[3 + k], __shfl_xor(v[3 + k], o, kWave));
    }
  int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 6; k++) red[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double a = red[k][0], b = red[3 + k][0];
      for (int q = 1; q < kBlock / kWave; q++) { a = dmin(a, red[k][q]); b = dmax(b, red[3 + k][q]); }
      part[6 * blockIdx.x + k] = a;
      part[6 * blockIdx.x + 3 + k] = b;
    }
  }
}

__device__ inline uint32_t expand_bits(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ void k_scene_final(int nPart, const double* __restrict__ part, double* __restrict__ scene) {
  __shared__ double red[6][kBlock / kWave];
  double v[6];
#pragma unroll
  for (int k = 0; k < 3; k++) { v[k] = 1.7976931348623157e308; v[3 + k] = -1.7976931348623157e308; }
  for (int q = threadIdx.x; q < nPart; q += blockDim.x)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      v[k] = dmin(v[k], part[6 * q + k]);
      v[3 + k] = dmax(v[3 + k], part[6 * q + 3 + k]);
    }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      v[k] = dmin(v[k], __shfl_xor(v[k], o, kWave));
      v[3 + k] = dmax(v[3 + k], __shfl_xor(v[3 + k], o, kWave));
    }
  int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 6; k++) red[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; k++) {
      double a = red[k][0], b = red[3 + k][0];
      for (int q = 1; q < kBlock / kWave; q++) { a = dmin(a, red[k][q]); b = dmax(b, red[3 + k][q]); }
      scene[k] = a;
      scene[3 + k] = b;
    }
}

// morton_code(double4) of lbvh/morton_code.cuh:34-44 applied to the box
// centroid normalised to the scene box (default_morton_code_calculator, bvh.cuh:197-221)
__global__ void k_morton(int P, const double* __restrict__ scene, const double* __restrict__ box,
                         uint32_t* __restrict__ code, int* __restrict__ idx) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  uint32_t e[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double c = (box[(3 + k) * P + p] + box[k * P + p]) * 0.5;
    c -= scene[k];
    c /= (scene[3 + k] - scene[k]);
    double r = fmin(fmax(c * 1024.0, 0.0), 1023.0);
    e[k] = expand_bits((uint32_t)r);
  }
  code[p] = e[0] * 4 + e[1] * 2 + e[2];
  idx[p] = p;
}

// ---------------- stable LSD radix sort (8-bit digits) ----------------
__global__ void k_radix_hist(int n, int shift, const uint32_t* __restrict__ keys, int* __restrict__ hist) {
  __shared__ int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  int base = blockIdx.x * kRadixTile;
#pragma unroll
  for (int j = 0; j < kRadixItems; j++) {
    int i = base + j * kBlock + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1);
  }
  __syncthreads();
  hist[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// keys of a tile are visited in order (round j covers tile positions j*256 + t);
// ranks among equal digits come from 8 wavefront ballots, then from the waves
// before this one, then from the earlier rounds -- a stable scatter.
__global__ void k_radix_scatter(int n, int shift, const uint32_t* __restrict__ keysIn, const int* __restrict__ idxIn,
                                const int* __restrict__ histScan, uint32_t* __restrict__ keysOut,
                                int* __restrict__ idxOut) {
  constexpr int NW = kBlock / kWave;
  __shared__ int waveCnt[NW][256];
  __shared__ int digitBase[256];
  __shared__ int globalBase[256];
  int t = threadIdx.x, lane = t % kWave, w = t / kWave;
  digitBase[t] = 0;
  globalBase[t] = histScan[t * gridDim.x + blockIdx.x];
#pragma unroll
  for (int q = 0; q < NW; q++) waveCnt[q][t] = 0;
  __syncthreads();
  int base = blockIdx.x * kRadixTile;
  const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  for (int j = 0; j < kRadixItems; j++) {
    int i = base + j * kBlock + t;
    bool valid = i < n;
    uint32_t key = valid ? keysIn[i] : 0u;
    int val = valid ? idxIn[i] : 0;
    int d = (int)((key >> shift) & 255u);
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      unsigned long long bb = __ballot((d >> b) & 1);
      peers &= ((d >> b) & 1) ? bb : ~bb;
    }
    int rank = __popcll(peers & lt);
    int leader = valid ? (__ffsll((long long)peers) - 1) : -1;
    if (valid && lane == leader) waveCnt[w][d] = __popcll(peers);
    __syncthreads();
    if (valid) {
      int pos = digitBase[d] + rank;
      for (int q = 0; q < w; q++) pos += waveCnt[q][d];
      int dst = globalBase[d] + pos;
      keysOut[dst] = key;
      idxOut[dst] = val;
    }
    __syncthreads();
    int add = 0;
#pragma unroll
    for (int q = 0; q < NW; q++) { add += waveCnt[q][t]; waveCnt[q][t] = 0; }
    digitBase[t] += add;
    __syncthreads();
  }
}

// ---------------- exclusive scan of int counts ----------------
__device__ inline int block_exclusive_scan(int v, int* sh, int* total) {
  // inclusive scan within waves via shuffles, then across waves
  int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  int x = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    int y = __shfl_up(x, o, kWave);
    if (lane >= o) x += y;
  }
  if (lane == kWave - 1) sh[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int q = 0; q < kBlock / kWave; q++) { int c = sh[q]; sh[q] = s; s += c; }
    sh[kBlock / kWave] = s;
  }
  __syncthreads();
  int r = x - v + sh[w];
  *total = sh[kBlock / kWave];
  __syncthreads();
  return r;
}

// Exclusive scan of at most kScanOneMax counts in one workgroup (one launch
// instead of three): the workgroup walks the array in coalesced tiles of
// kScanOneBlock * kScanItems counts, carrying the running total.  In-place use
// (in == out) is allowed: every count is read before its slot is written.
constexpr int kScanOneBlock = 1024;
constexpr int kScanOneMax = 1 << 14;
__global__ void __launch_bounds__(kScanOneBlock) k_scan_one(int n, const int* in, int* out, int* __restrict__ total) {
  __shared__ int sh[kScanOneBlock / kWave];
  __shared__ int tileTotal;
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  int carry = 0;
  for (int base = 0; base < n; base += kScanOneBlock * kScanItems) {
    const int i0 = base + threadIdx.x * kScanItems;
    int v[kScanItems], s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
      v[k] = i0 + k < n ? in[i0 + k] : 0;
      s += v[k];
    }
    int x = s;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      int y = __shfl_up(x, o, kWave);
      if (lane >= o) x += y;
    }
    if (lane == kWave - 1) sh[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      int c = 0;
      for (int q = 0; q < kScanOneBlock / kWave; q++) { int t = sh[q]; sh[q] = c; c += t; }
    }
    __syncthreads();
    int pre = carry + x - s + sh[w];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
      if (i0 + k < n) out[i0 + k] = pre;
      pre += v[k];
    }
    if (threadIdx.x == kScanOneBlock - 1) tileTotal = pre - carry;  // the last thread's inclusive prefix
    __syncthreads();
    carry += tileTotal;
    __syncthreads();
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void k_scan_tiles(int n, const int* __restrict__ in, int* __restrict__ out, int* __restrict__ blockSums) {
  __shared__ int sh[kBlock / kWave + 1];
  int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems], s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  int tot;
  int pre = block_exclusive_scan(s, sh, &tot);
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    if (base + k < n) out[base + k] = pre;
    pre += v[k];
  }
  if (threadIdx.x == 0) blockSums[blockIdx.x] = tot;
}

__global__ void k_scan_sums(int nb, int* __restrict__ sums) {
  __shared__ int sh[kBlock / kWave + 1];
  int carry = 0;
  for (int base = 0; base < nb; base += kScanTile) {
    int idx = base + threadIdx.x * kScanItems;
    int v[kScanItems], s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
      v[k] = (idx + k < nb) ? sums[idx + k] : 0;
      s += v[k];
    }
    int tot;
    int pre = block_exclusive_scan(s, sh, &tot) + carry;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
      if (idx + k < nb) sums[idx + k] = pre;
      pre += v[k];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) sums[nb] = carry;
}

__global__ void k_scan_add(int n, int nb, const int* __restrict__ sums, int* __restrict__ out, int* __restrict__ total) {
  int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int add = sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; k++)
    if (base + k < n) out[base + k] += add;
  if (blockIdx.x == 0 && threadIdx.x == 0 && total) *total = sums[nb];
}

// k_scan_add with the tile prefix summed in place: workgroup b adds up the
// first b tile sums itself (at most kScanDirectMax of them), so a scan of up
// to kScanDirectMax tiles is two launches instead of three.
constexpr int kScanDirectMax = 1024;
__global__ void k_scan_add_direct(int n, int nb, const int* __restrict__ blockSums, int* __restrict__ out,
                                  int* __restrict__ total) {
  __shared__ int sh[kBlock / kWave];
  const int last = blockIdx.x == nb - 1 ? nb : (int)blockIdx.x;  // the last workgroup also sums its own: the total
  int acc = 0;
  for (int i = threadIdx.x; i < last; i += kBlock) acc += blockSums[i];
  for (int o = kWave / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
  if (threadIdx.x % kWave == 0) sh[threadIdx.x / kWave] = acc;
  __syncthreads();
  int sum = 0;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; w++) sum += sh[w];
  const int add = blockIdx.x == nb - 1 ? sum - blockSums[nb - 1] : sum;
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
#pragma unroll
  for (int k = 0; k < kScanItems; k++)
    if (base + k < n) out[base + k] += add;
  if (blockIdx.x == nb - 1 && threadIdx.x == 0 && total) *total = sum;
}

// ---------------- Karras LBVH ----------------
__device__ inline int cub64(uint64_t a, uint64_t b) { return __clzll((long long)(a ^ b)); }

__global__ void k_make_keys(int P, const uint32_t* __restrict__ code, const int* __restrict__ idx,
                            uint64_t* __restrict__ keys) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < P) keys[i] = ((uint64_t)code[i] << 32) | (uint32_t)idx[i];
}

__global__ void k_karras(int P, const uint64_t* __restrict__ keys, int* __restrict__ left, int* __restrict__ right,
                         int* __restrict__ parent, int2* __restrict__ range) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= P - 1) return;
  int first, last;
  if (idx == 0) {
    first = 0;
    last = P - 1;
  } else {
    uint64_t self = keys[idx];
    int L = cub64(self, keys[idx - 1]);
    int R = cub64(self, keys[idx + 1]);
    int d = R > L ? 1 : -1;
    int dmin_ = L < R ? L : R;
    int lmax = 2, delta = -1;
    int it = idx + d * lmax;
    if (0 <= it && it < P) delta = cub64(self, keys[it]);
    while (delta > dmin_) {
      lmax <<= 1;
      it = idx + d * lmax;
      delta = -1;
      if (0 <= it && it < P) delta = cub64(self, keys[it]);
    }
    int l = 0, t = lmax >> 1;
    while (t > 0) {
      it = idx + (l + t) * d;
      delta = -1;
      if (0 <= it && it < P) delta = cub64(self, keys[it]);
      if (delta > dmin_) l += t;
      t >>= 1;
    }
    int jdx = idx + l * d;
    first = d < 0 ? jdx : idx;
    last = d < 0 ? idx : jdx;
  }
  // find_split
  uint64_t fc = keys[first], lc = keys[last];
  int gamma;
  if (fc == lc) {
    gamma = (first + last) >> 1;
  } else {
    int dn = cub64(fc, lc);
    int split = first, stride = last - first;
    do {
      stride = (stride + 1) >> 1;
      int middle = split + stride;
      if (middle < last) {
        int dd = cub64(fc, keys[middle]);
        if (dd > dn) split = middle;
      }
    } while (stride > 1);
    gamma = split;
  }
  int lc_ = gamma, rc_ = gamma + 1;
  if (first == gamma) lc_ += P - 1;
  if (last == gamma + 1) rc_ += P - 1;
  left[idx] = lc_;
  right[idx] = rc_;
  range[idx] = make_int2(first, last);
  parent[lc_] = idx;
  parent[rc_] = idx;
  if (idx == 0) parent[0] = -1;
}

// Node boxes without inter-block synchronisation: every internal node covers a
// contiguous leaf range [first, last] (Karras), so its box is the union of the
// leaf boxes over that range, read from a sparse table of power-of-two window
// unions (level j holds the union of leaves [i, i + 2^j)).  Min / max are exact
// and order-free, so the result equals the bottom-up refit bit for bit.
__global__ void k_table_level0(int P, const int* __restrict__ sortedIdx, const double* __restrict__ pbox,
                               double* __restrict__ tab) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  int obj = sortedIdx[i];
#pragma unroll
  for (int k = 0; k < 6; k++) tab[(int64_t)k * P + i] = pbox[(int64_t)k * P + obj];
}

// Levels j0 .. j0 + count - 1 of the sparse table in one launch, each from
// level j0 - 1: level L at i is the union of the leaf boxes [i, i + 2^L), i.e.
// of 2^(L - j0 + 1) consecutive entries of level j0 - 1 at stride 2^(j0 - 1),
// folded in order (min / max are exact, so every level equals the pairwise
// construction's).
// levels per launch: 4 while the table is small (the launches, not the
// reads, set its cost), 2 beyond (the level reads then cost as in the
// pairwise construction: 4 entries per output position for 2 levels)
inline int table_levels_per_launch(int P) { return P < (1 << 17) ? 4 : 2; }
__global__ void k_table_levels(int P, int j0, int count, double* __restrict__ tab) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int half = 1 << (j0 - 1);
  if (i + 2 * half > P) return;
  const double* prev = tab + (int64_t)(j0 - 1) * 6 * P;
  double lo[3], hi[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lo[k] = prev[(int64_t)k * P + i];
    hi[k] = prev[(int64_t)(3 + k) * P + i];
  }
  int t = 1;  // entries of level j0 - 1 folded so far
  for (int m = 0; m < count; m++) {
    const int need = 2 << m;
    if (i + (int64_t)need * half > P) break;
    for (; t < need; t++)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        lo[k] = dmin(lo[k], prev[(int64_t)k * P + i + t * half]);
        hi[k] = dmax(hi[k], prev[(int64_t)(3 + k) * P + i + t * half]);
      }
    double* cur = tab + (int64_t)(j0 + m) * 6 * P;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      cur[(int64_t)k * P + i] = lo[k];
      cur[(int64_t)(3 + k) * P + i] = hi[k];
    }
  }
}

// Node boxes rounded outward to float, child links and preorder skip links.
// The skip link of node n is the right sibling of the nearest ancestor-or-self
// that is a left child (-1 when there is none: n ends the root's preorder).
// Internal nodes over at most kSubLeaves leaves carry kSmall in their child
// link: the top-level traversal hands such subtrees to the second pass.
constexpr int kSubLeaves = 32;
constexpr int kSmall = 1 << 30;
constexpr int kChildMask = kSmall - 1;

__global__ void k_node_boxes(int P, const int2* __restrict__ range, const double* __restrict__ tab,
                             const int* __restrict__ left, const int* __restrict__ right,
                             const int* __restrict__ parent, const int* __restrict__ sortedIdx,
                             BvhNode* __restrict__ nodes) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int NN = 2 * P - 1;
  if (n >= NN) return;
  double lo[3], hi[3];
  BvhNode nd;
  if (n >= P - 1) {
    int i = n - (P - 1);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      lo[k] = tab[(int64_t)k * P + i];
      hi[k] = tab[(int64_t)(3 + k) * P + i];
    }
    nd.child = sortedIdx[i];
  } else {
    int2 r = range[n];
    int len = r.y - r.x + 1;
    int j = 31 - __clz(len);
    const double* t = tab + (int64_t)j * 6 * P;
    int a = r.x, b = r.y - (1 << j) + 1;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      lo[k] = dmin(t[(int64_t)k * P + a], t[(int64_t)k * P + b]);
      hi[k] = dmax(t[(int64_t)(3 + k) * P + a], t[(int64_t)(3 + k) * P + b]);
    }
    nd.child = left[n] | (len <= kSubLeaves ? kSmall : 0);
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    nd.lo[k] = __double2float_rd(lo[k]);
    nd.hi[k] = __double2float_ru(hi[k]);
  }
  int skip = -1, cur = n;
  while (cur != 0) {  // node 0 is the root (the only leaf when P == 1)
    int pa = parent[cur];
    if (left[pa] == cur) {
      skip = right[pa];
      break;
    }
    cur = pa;
  }
  nd.skip = skip;
  nodes[n] = nd;
}

__device__ inline bool overlap(const double* a, const double* b) {
  if (a[3] < b[0] || b[3] < a[0]) return false;
  if (a[4] < b[1] || b[4] < a[1]) return false;
  if (a[5] < b[2] || b[5] < a[2]) return false;
  return true;
}

// Append to a wavefront-owned segment: lanes with `flag` get consecutive slots
// after `count` (the wavefront's running count, the same in all its lanes that
// are still looping), in lane order; nothing past the segment's capacity is
// written, but the count keeps counting (the host sees the overflow, grows the
// segments and repeats).  No atomics: a wavefront's output depends only on its
// own inputs, so the order is reproducible.
__device__ inline void seg_append(bool flag, int2 v, int& count, int2* __restrict__ seg, int cap) {
  unsigned long long m = __ballot(flag);
  if (!m) return;
  int lane = threadIdx.x % kWave;
  const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  if (flag) {
    int pos = count + (int)__popcll(m & lt);
    if (pos < cap) seg[pos] = v;
  }
  count += (int)__popcll(m);
}

struct TraverseArgs {
  int P;
  const double* pbox;
  const BvhNode* nodes;
  const uint8_t* patchStatic;
  int skipStatic;
};

__device__ inline void load_query(const TraverseArgs& a, int p, double q[6], float qlo[3], float qhi[3]) {
#pragma unroll
  for (int k = 0; k < 6; k++) q[k] = a.pbox[(int64_t)k * a.P + p];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    qlo[k] = __double2float_rd(q[k]);
    qhi[k] = __double2float_ru(q[3 + k]);
  }
}

__device__ inline bool overlap_f(const float qlo[3], const float qhi[3], const BvhNode& nd) {
  return !(qhi[0] < nd.lo[0] || nd.hi[0] < qlo[0] || qhi[1] < nd.lo[1] || nd.hi[1] < qlo[1] || qhi[2] < nd.lo[2] ||
           nd.hi[2] < qlo[2]);
}

// a leaf reached with an overlapping float box: the exact FP64 test and the
// emission rule -- every pair p < q once: by p, or, when static queries are
// skipped, by the simulated side of a static/simulated pair
__device__ inline bool leaf_hit(const TraverseArgs& a, int p, const double q[6], bool ov, int obj) {
  if (!ov || !(obj > p || (a.skipStatic && obj < p && a.patchStatic[obj]))) return false;
  double b[6];
#pragma unroll
  for (int k = 0; k < 6; k++) b[k] = a.pbox[(int64_t)k * a.P + obj];
  return overlap(q, b);
}

// Two-pass stackless traversal (latency-bound walks are kept short and
// numerous).  Pass 1: lane t takes the patch at sorted (Morton) position t,
// walks the top of the tree and, instead of descending into a subtree of at
// most kSubLeaves leaves whose float box overlaps, emits the work item
// (patch, subtree root).  Pass 2: one lane per item walks that subtree, which
// yields at most kSubLeaves pairs.  Internal nodes are pruned on their
// outward-rounded float boxes; a leaf is confirmed on the exact FP64 boxes
// (closed-box test), so the pair set is exactly that of the FP64 boxes.
// Static patches follow all simulated ones in patch order, so with skipStatic
// a static query has nothing to report.  Each wavefront writes segments of its
// own; their lengths go to cntI / cntP (one entry per wavefront).
__global__ void __launch_bounds__(kBlock) k_traverse_top(TraverseArgs a, const int* __restrict__ sortedIdx,
                                                         int2* __restrict__ segI, int capI, int* __restrict__ cntI,
                                                         int2* __restrict__ segP, int capP, int* __restrict__ cntP,
                                                         int* __restrict__ overflow) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int w = t / kWave;
  int p = t < a.P ? sortedIdx[t] : 0;
  bool alive = t < a.P && !(a.skipStatic && a.patchStatic[p]);
  double q[6] = {0, 0, 0, 0, 0, 0};
  float qlo[3] = {0, 0, 0}, qhi[3] = {0, 0, 0};
  if (alive) load_query(a, p, q, qlo, qhi);
  int nI = 0, nP = 0;
  int2* sI = segI + (int64_t)w * capI;
  int2* sP = segP + (int64_t)w * capP;
  int node = alive ? 0 : -1;
  while (node >= 0) {
    BvhNode nd = a.nodes[node];
    bool ov = overlap_f(qlo, qhi, nd);
    bool hit = false, item = false;
    int obj = -1;
    if (node >= a.P - 1) {
      obj = nd.child;
      hit = leaf_hit(a, p, q, ov, obj);
      node = nd.skip;
    } else if (ov && (nd.child & kSmall)) {
      item = true;
      obj = node;
      node = nd.skip;
    } else {
      node = ov ? (nd.child & kChildMask) : nd.skip;
    }
    seg_append(hit, make_int2(min(p, obj), max(p, obj)), nP, sP, capP);
    seg_append(item, make_int2(p, obj), nI, sI, capI);
  }
  // the running counts grow monotonically and the last lanes to leave the loop
  // hold the totals: take the maximum over the wavefront
  for (int o = kWave / 2; o > 0; o >>= 1) {
    nI = max(nI, __shfl_xor(nI, o, kWave));
    nP = max(nP, __shfl_xor(nP, o, kWave));
  }
  if ((threadIdx.x % kWave) == 0) {
    cntI[w] = nI;
    cntP[w] = nP;
    if (nI > capI || nP > capP) atomicMax(overflow, max(nI, nP));
  }
}

// one lane per item; item i's pairs (at most kSubLeaves) go to its own segment
// seg[i * kSubLeaves ...], so the pair order follows the item order
__global__ void __launch_bounds__(kBlock) k_traverse_sub(TraverseArgs a, int nItems, const int2* __restrict__ items,
                                                         int2* __restrict__ seg, int* __restrict__ cnt) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nItems) return;
  int2 it = items[t];
  int p = it.x;
  double q[6];
  float qlo[3], qhi[3];
  load_query(a, p, q, qlo, qhi);
  int n = 0;
  int2* sg = seg + (int64_t)t * kSubLeaves;
  int node = it.y;
  const int stop = a.nodes[node].skip;
  while (node != stop) {
    BvhNode nd = a.nodes[node];
    bool ov = overlap_f(qlo, qhi, nd);
    if (node >= a.P - 1) {
      int obj = nd.child;
      if (leaf_hit(a, p, q, ov, obj) && n < kSubLeaves) sg[n++] = make_int2(min(p, obj), max(p, obj));
      node = nd.skip;
    } else {
      node = ov ? (nd.child & kChildMask) : nd.skip;
    }
  }
  cnt[t] = n;
}

// entry k of item i's segment goes to out[off[i] + k] when k < cnt[i]
__global__ void k_compact_items(int nItems, const int2* __restrict__ seg, const int* __restrict__ cnt,
                                const int* __restrict__ off, int2* __restrict__ out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nItems * kSubLeaves) return;
  int i = (int)(e / kSubLeaves), k = (int)(e % kSubLeaves);
  if (k < cnt[i]) out[off[i] + k] = seg[e];
}

// segment w's first min(cnt[w], cap) entries go to out[off[w] ...] (off: exclusive scan of cnt)
__global__ void k_compact(int nw, int cap, const int2* __restrict__ seg, const int* __restrict__ cnt,
                          const int* __restrict__ off, int2* __restrict__ out) {
  int w = blockIdx.x;
  if (w >= nw) return;
  int n = min(cnt[w], cap), o = off[w];
  for (int k = threadIdx.x; k < n; k += blockDim.x) out[o + k] = seg[(int64_t)w * cap + k];
}

// Per pair: point-triangle records (points of A x faces of B, then points of
// B x faces of A) into cntPT, edge-edge records (edges of A x edges of B)
// into cntEE.  All point-triangle records precede all edge-edge ones, so the
// wavefronts of the narrow-phase kernels see one record type.
__global__ void k_rec_count(int Q, const int2* __restrict__ pairs, const int* __restrict__ ppOff,
                            const int* __restrict__ pfOff, const int* __restrict__ peOff, int* __restrict__ cntPT,
                            int* __restrict__ cntEE) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Q) return;
  int2 pr = pairs[i];
  int pa = ppOff[pr.x + 1] - ppOff[pr.x], fa = pfOff[pr.x + 1] - pfOff[pr.x], ea = peOff[pr.x + 1] - peOff[pr.x];
  int pb = ppOff[pr.y + 1] - ppOff[pr.y], fb = pfOff[pr.y + 1] - pfOff[pr.y], eb = peOff[pr.y + 1] - peOff[pr.y];
  cntPT[i] = pa * fb + pb * fa;
  cntEE[i] = ea * eb;
}

// Both record-fill kernels write the same records in the same slots: one lane
// per record while the pairs are few (456 K pairs, 5.5 M records for a 256^2
// cloth over a subdivision-6 icosphere: 567 us per broad phase against 580
// with one lane per pair, too few lanes to hide the pair's list loads), one
// lane per pair from kRecFillPairsMin pairs (18 M pairs, 265 M records at the
// 3 x 1024^2 scale: 22.0 -> 19.2 ms per broad phase, each pair's offsets and
// lists read once instead of per record; the 256^2 scene forced to one lane
// per pair gave bit-identical steps).
constexpr int kRecFillPairsMin = 1 << 21;

__device__ inline int last_le(const int* __restrict__ off, int lo, int hi, int r) {
  while (lo < hi) {  // last i in [lo, hi] with off[i] <= r
    int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= r) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// One lane per record (below kRecFillPairsMin patch pairs): records [0, nPT)
// are point-triangle ones, the pair
// found in offPT; records [nPT, K) edge-edge ones, the pair found in offEE.
// Within a pair the order is the reference's (getCcdPotentialICPairsListKernel,
// PdIpcGpuFunc.cu:1838-1948) restricted to the record type.  A block whose
// records lie in one segment and span at most kBlock pairs searches a copy
// of its offsets in LDS.
__global__ void __launch_bounds__(kBlock) k_rec_fill_records(int Q, int nPT, int K, const int2* __restrict__ pairs,
                                                     const int* __restrict__ offPT, const int* __restrict__ offEE,
                                                     const int* __restrict__ ppOff, const int* __restrict__ ppList,
                                                     const int* __restrict__ pfOff, const int* __restrict__ pfList,
                                                     const int* __restrict__ peOff, const int* __restrict__ peList,
                                                     const int3* __restrict__ faces, const int2* __restrict__ edges,
                                                     int4* __restrict__ recs) {
  __shared__ int sRange[2];
  __shared__ int sOff[kBlock + 1];
  int r0 = blockIdx.x * kBlock;
  int r1 = min(r0 + kBlock, K) - 1;
  bool ee0 = r0 >= nPT, oneSeg = ee0 == (r1 >= nPT);
  const int* off0 = ee0 ? offEE : offPT;
  int base0 = ee0 ? nPT : 0;
  if (oneSeg && threadIdx.x == 0) sRange[0] = last_le(off0, 0, Q - 1, r0 - base0);
  if (oneSeg && threadIdx.x == 1) sRange[1] = last_le(off0, 0, Q - 1, r1 - base0);
  __syncthreads();
  int lo = oneSeg ? sRange[0] : 0, hi = oneSeg ? sRange[1] : 0;
  bool local = oneSeg && hi - lo + 1 <= kBlock;
  if (local && threadIdx.x <= hi - lo) sOff[threadIdx.x] = off0[lo + threadIdx.x];
  __syncthreads();
  int r = r0 + threadIdx.x;
  if (r >= K) return;
  bool ee = r >= nPT;
  const int* off = ee ? offEE : offPT;
  int rr = ee ? r - nPT : r;
  int pi = local ? lo + last_le(sOff, 0, hi - lo, rr) : last_le(off, 0, Q - 1, rr);
  int2 pr = pairs[pi];
  int id = rr - off[pi];
  int A = pr.x, B = pr.y;
  int4 out;
  if (!ee) {
    int pa = ppOff[A + 1] - ppOff[A], fa = pfOff[A + 1] - pfOff[A];
    int fb = pfOff[B + 1] - pfOff[B];
    int VaFb = pa * fb;
    if (id < VaFb) {
      int3 f = faces[pfList[pfOff[B] + id % fb]];
      out = make_int4(ppList[ppOff[A] + id / fb], f.x, f.y, f.z);
    } else {
      id -= VaFb;
      int3 f = faces[pfList[pfOff[A] + id % fa]];
      out = make_int4(ppList[ppOff[B] + id / fa], f.x, f.y, f.z);
    }
  } else {
    int eb = peOff[B + 1] - peOff[B];
    int2 e1 = edges[peList[peOff[A] + id / eb]];
    int2 e2 = edges[peList[peOff[B] + id % eb]];
    out = make_int4(~e1.x, e1.y, e2.x, e2.y);
  }
  recs[r] = out;
}

// One lane per patch pair (from kRecFillPairsMin pairs) writes all of the
// pair's records: the point-triangle ones from offPT[i], the edge-edge ones
// from nPT + offEE[i], each in the
// reference's order within the pair (getCcdPotentialICPairsListKernel,
// PdIpcGpuFunc.cu:1838-1948) restricted to the record type: A's points x B's
// faces, then B's points x A's faces; A's edges x B's edges.  The pair's
// offsets and lists are read once for all its records (about 13 on the
// cloth), instead of a search and the same loads per record.
__global__ void __launch_bounds__(kBlock) k_rec_fill_pairs(int Q, int nPT, const int2* __restrict__ pairs,
                                                     const int* __restrict__ offPT, const int* __restrict__ offEE,
                                                     const int* __restrict__ ppOff, const int* __restrict__ ppList,
                                                     const int* __restrict__ pfOff, const int* __restrict__ pfList,
                                                     const int* __restrict__ peOff, const int* __restrict__ peList,
                                                     const int3* __restrict__ faces, const int2* __restrict__ edges,
                                                     int4* __restrict__ recs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Q) return;
  const int2 pr = pairs[i];
  const int A = pr.x, B = pr.y;
  const int pA = ppOff[A], nPA = ppOff[A + 1] - pA, pB = ppOff[B], nPB = ppOff[B + 1] - pB;
  const int fA = pfOff[A], nFA = pfOff[A + 1] - fA, fB = pfOff[B], nFB = pfOff[B + 1] - fB;
  const int eA = peOff[A], nEA = peOff[A + 1] - eA, eB = peOff[B], nEB = peOff[B + 1] - eB;
  int o = offPT[i];
  for (int u = 0; u < nPA; u++) {
    const int pt = ppList[pA + u];
    for (int f = 0; f < nFB; f++) {
      const int3 t = faces[pfList[fB + f]];
      recs[o++] = make_int4(pt, t.x, t.y, t.z);
    }
  }
  for (int u = 0; u < nPB; u++) {
    const int pt = ppList[pB + u];
    for (int f = 0; f < nFA; f++) {
      const int3 t = faces[pfList[fA + f]];
      recs[o++] = make_int4(pt, t.x, t.y, t.z);
    }
  }
  int q = nPT + offEE[i];
  for (int a = 0; a < nEA; a++) {
    const int2 e1 = edges[peList[eA + a]];
    for (int b = 0; b < nEB; b++) {
      const int2 e2 = edges[peList[eB + b]];
      recs[q++] = make_int4(~e1.x, e1.y, e2.x, e2.y);
    }
  }
}

// ---------------- zero-intersection check (validation, no counterpart in the reference) ----------------
// A triangle pair intersects (non-coplanar) iff an edge of one crosses the
// other, so the check counts edge-triangle crossings, in FP64, over every
// (edge, face) of overlapping patch pairs (both ways) and within each patch,
// skipping pairs that share a vertex.  orient3 > 0: d above the plane (a, b, c).
__device__ inline double orient3(const double* a, const double* b, const double* c, const double* d) {
  double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  double v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  double w[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
  return u[0] * (v[1] * w[2] - v[2] * w[1]) + u[1] * (v[2] * w[0] - v[0] * w[2]) + u[2] * (v[0] * w[1] - v[1] * w[0]);
}

// the segment (p, q) meets the triangle (a, b, c) in one point (coplanar
// configurations do not count)
__device__ inline bool edge_crosses(const double* p, const double* q, const double* a, const double* b,
                                    const double* c) {
  double s1 = orient3(a, b, c, p), s2 = orient3(a, b, c, q);
  if ((s1 > 0 && s2 > 0) || (s1 < 0 && s2 < 0) || (s1 == 0 && s2 == 0)) return false;
  double t1 = orient3(p, q, a, b), t2 = orient3(p, q, b, c), t3 = orient3(p, q, c, a);
  bool pos = t1 >= 0 && t2 >= 0 && t3 >= 0, neg = t1 <= 0 && t2 <= 0 && t3 <= 0;
  return (pos || neg) && !(t1 == 0 && t2 == 0 && t3 == 0);
}

// blocks [0, Q): patch pairs; blocks [Q, Q + P): single patches; one wavefront each
__global__ void __launch_bounds__(kWave) k_count_crossings(int Q, int P, int ld, const int2* __restrict__ pairs,
                                                           const int* __restrict__ pfOff, const int* __restrict__ pfList,
                                                           const int* __restrict__ peOff, const int* __restrict__ peList,
                                                           const int3* __restrict__ faces,
                                                           const int2* __restrict__ edges, const double* __restrict__ X,
                                                           unsigned long long* __restrict__ count) {
  int b = blockIdx.x;
  int A, B;
  if (b < Q) {
    int2 pr = pairs[b];
    A = pr.x;
    B = pr.y;
  } else {
    A = B = b - Q;
    if (A >= P) return;
  }
  int ea = peOff[A + 1] - peOff[A], fa = pfOff[A + 1] - pfOff[A];
  int eb = peOff[B + 1] - peOff[B], fb = pfOff[B + 1] - pfOff[B];
  int n1 = ea * fb, n = A == B ? n1 : n1 + eb * fa;
  unsigned long long hits = 0;
  for (int c = threadIdx.x; c < n; c += kWave) {
    int e, f;
    if (c < n1) {
      e = peList[peOff[A] + c / fb];
      f = pfList[pfOff[B] + c % fb];
    } else {
      int d = c - n1;
      e = peList[peOff[B] + d / fa];
      f = pfList[pfOff[A] + d % fa];
    }
    int2 ed = edges[e];
    int3 fc = faces[f];
    if (ed.x == fc.x || ed.x == fc.y || ed.x == fc.z || ed.y == fc.x || ed.y == fc.y || ed.y == fc.z) continue;
    double p[3], q[3], a[3], bb[3], cc[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      p[k] = X[k * ld + ed.x];
      q[k] = X[k * ld + ed.y];
      a[k] = X[k * ld + fc.x];
      bb[k] = X[k * ld + fc.y];
      cc[k] = X[k * ld + fc.z];
    }
    if (edge_crosses(p, q, a, bb, cc)) hits++;
  }
  for (int o = kWave / 2; o > 0; o >>= 1) hits += __shfl_xor(hits, o, kWave);
  if (threadIdx.x == 0 && hits) atomicAdd(count, hits);
}

}  // namespace

void launch_exclusive_scan(int* in, int* out, int n, int* total_out, int* blockScratch, hipStream_t s) {
  if (n <= 0) return;
  if (n <= kScanOneMax) {
    hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(kScanOneBlock), 0, s, n, in, out, total_out);
    return;
  }
  int nb = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_tiles, dim3(nb), dim3(kBlock), 0, s, n, in, out, blockScratch);
  if (nb <= kScanDirectMax) {
    hipLaunchKernelGGL(k_scan_add_direct, dim3(nb), dim3(kBlock), 0, s, n, nb, blockScratch, out, total_out);
    return;
  }
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, s, nb, blockScratch);
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kBlock), 0, s, n, nb, blockScratch, out, total_out);
}

void launch_patch_boxes(const DeviceMesh& m, const double* x, const double* dx, double inflation, hipStream_t s) {
  int P = m.nPatches;
  if (P <= 0) return;
  hipLaunchKernelGGL(k_patch_boxes, dim3(((int64_t)P * kBoxLanes + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P,
                     m.ld, m.pfOff, m.pfList, m.bFaces, x, dx, m.nSim, inflation, m.patchBox);
}

void launch_lbvh(DeviceMesh& m, hipStream_t s) {
  int P = m.nPatches;
  if (P <= 0) return;
  int nPart = std::min(256, (P + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_scene_partial, dim3(nPart), dim3(kBlock), 0, s, P, m.patchBox, m.sceneBoxPart);
  hipLaunchKernelGGL(k_scene_final, dim3(1), dim3(kBlock), 0, s, nPart, m.sceneBoxPart, m.sceneBox);
  hipLaunchKernelGGL(k_morton, dim3((P + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P, m.sceneBox, m.patchBox,
                     m.mortonA, m.idxA);
  int nb = (P + kRadixTile - 1) / kRadixTile;
  uint32_t *kin = m.mortonA, *kout = m.mortonB;
  int *iin = m.idxA, *iout = m.idxB;
  for (int shift = 0; shift < 30; shift += 8) {
    hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kBlock), 0, s, P, shift, kin, m.radixHist);
    launch_exclusive_scan(m.radixHist, m.radixHist, 256 * nb, nullptr, m.radixBlockSums, s);
    hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kBlock), 0, s, P, shift, kin, iin, m.radixHist, kout, iout);
    std::swap(kin, kout);
    std::swap(iin, iout);
  }
  // 4 passes: the sorted data is back in the A buffers
  hipLaunchKernelGGL(k_make_keys, dim3((P + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P, kin, iin, m.keys);
  if (P > 1)
    hipLaunchKernelGGL(k_karras, dim3((P - 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P, m.keys, m.nodeLeft,
                       m.nodeRight, m.nodeParent, m.nodeRange);
  hipLaunchKernelGGL(k_table_level0, dim3((P + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P, iin, m.patchBox,
                     m.nodeTable);
  int top = 0;  // the highest level: 2^top <= P
  while ((2 << top) <= P) top++;
  const int per = table_levels_per_launch(P);
  for (int j = 1; j <= top; j += per)
    hipLaunchKernelGGL(k_table_levels, dim3((P + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P, j,
                       std::min(per, top - j + 1), m.nodeTable);
  hipLaunchKernelGGL(k_node_boxes, dim3((2 * P - 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P, m.nodeRange,
                     m.nodeTable, m.nodeLeft, m.nodeRight, m.nodeParent, iin, m.nodes);
}

int traverse_top_waves(int P) { return ((P + kBlock - 1) / kBlock) * (kBlock / kWave); }
int traverse_sub_cap() { return kSubLeaves; }

void launch_traverse_top(const DeviceMesh& m, int skipStatic, hipStream_t s) {
  int P = m.nPatches;
  if (P <= 0) return;
  const TraverseArgs a{P, m.patchBox, m.nodes, m.patchStatic, skipStatic};
  int nw = traverse_top_waves(P);
  hipLaunchKernelGGL(k_traverse_top, dim3((P + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a, m.idxA, m.segI, m.segCapI,
                     m.cntI, m.segP, m.segCapP, m.cntP, m.counters + 3);
  launch_exclusive_scan(m.cntI, m.offI, nw, m.counters + 6, m.scanBlock, s);
  launch_exclusive_scan(m.cntP, m.offP, nw, m.counters + 7, m.scanBlock, s);
}

void launch_traverse_sub(const DeviceMesh& m, int skipStatic, int nItems, hipStream_t s) {
  if (nItems <= 0) return;
  int P = m.nPatches;
  int nw = traverse_top_waves(P);
  hipLaunchKernelGGL(k_compact, dim3(nw), dim3(kBlock), 0, s, nw, m.segCapI, m.segI, m.cntI, m.offI, m.items);
  const TraverseArgs a{P, m.patchBox, m.nodes, m.patchStatic, skipStatic};
  hipLaunchKernelGGL(k_traverse_sub, dim3((nItems + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a, nItems, m.items,
                     m.segS, m.cntS);
  launch_exclusive_scan(m.cntS, m.offS, nItems, m.counters + 0, m.scanBlock, s);
}

void launch_pairs_compact(const DeviceMesh& m, int nItems, int nTopPairs, hipStream_t s) {
  int nw = traverse_top_waves(m.nPatches);
  if (nTopPairs > 0)
    hipLaunchKernelGGL(k_compact, dim3(nw), dim3(kBlock), 0, s, nw, m.segCapP, m.segP, m.cntP, m.offP, m.pairs);
  if (nItems > 0) {
    int64_t n = (int64_t)nItems * kSubLeaves;
    hipLaunchKernelGGL(k_compact_items, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, nItems,
                       m.segS, m.cntS, m.offS, m.pairs + nTopPairs);
  }
}

void launch_count_crossings(const DeviceMesh& m, int nPairs, unsigned long long* count, hipStream_t s) {
  int P = m.nPatches;
  if (P <= 0) return;
  hipLaunchKernelGGL(k_count_crossings, dim3((unsigned)(nPairs + P)), dim3(kWave), 0, s, nPairs, P, m.ld, m.pairs,
                     m.pfOff, m.pfList, m.peOff, m.peList, m.bFaces, m.bEdges, m.X, count);
}

void launch_rec_count(const DeviceMesh& m, int nPairs, hipStream_t s) {
  if (nPairs <= 0) return;
  hipLaunchKernelGGL(k_rec_count, dim3((nPairs + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nPairs, m.pairs, m.ppOff,
                     m.pfOff, m.peOff, m.recCount, m.recCount + m.pairCap);
  launch_exclusive_scan(m.recCount, m.recOff, nPairs, m.counters + 1, m.scanBlock, s);
  launch_exclusive_scan(m.recCount + m.pairCap, m.recOff + m.pairCap, nPairs, m.counters + 8, m.scanBlock, s);
}

void launch_rec_fill(const DeviceMesh& m, int nPairs, int nPT, int nRecs, hipStream_t s) {
  if (nRecs <= 0 || nPairs <= 0) return;
  if (nPairs >= kRecFillPairsMin)
    hipLaunchKernelGGL(k_rec_fill_pairs, dim3((nPairs + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nPairs, nPT,
                       m.pairs, m.recOff, m.recOff + m.pairCap, m.ppOff, m.ppList, m.pfOff, m.pfList, m.peOff,
                       m.peList, m.bFaces, m.bEdges, m.recs);
  else
    hipLaunchKernelGGL(k_rec_fill_records, dim3((nRecs + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nPairs, nPT,
                       nRecs, m.pairs, m.recOff, m.recOff + m.pairCap, m.ppOff, m.ppList, m.pfOff, m.pfList, m.peOff,
                       m.peList, m.bFaces, m.bEdges, m.recs);
}

}  // namespace pdipc
