// Device-wide exclusive prefix sums (out[i] = sum in[0..i), out[n] = total),
// used for the cell-slot offsets and the contact/incidence offsets of the
// collision pipeline (the reference's thrust::inclusive_scan calls,
// CollisionDetector.cu:586,605,614,650 and System.cu:1357).
//
// Three passes over 1024-element tiles: tile sums, a scan of the tile sums
// (recursively), and the tile-local scans with their offsets added.  Each
// 256-thread block scans its tile with 64-wide wavefront shuffles.
#include "dv_internal.h"

namespace dv {
namespace {

constexpr int kThreads = 256;
constexpr int kPer = 4;
constexpr int kTile = kThreads * kPer;

template <typename T>
__device__ T blockExclusiveScan(T v, T* total) {
  __shared__ T waveSums[kThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) waveSums[wave] = x;
  __syncthreads();
  T waveOff = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kThreads / 64; ++w) {
    if (w < wave) waveOff += waveSums[w];
    tot += waveSums[w];
  }
  __syncthreads();
  *total = tot;
  return x - v + waveOff;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void tileSums(const T* __restrict__ in, T* __restrict__ sums, int n) {
  const long base = (long)blockIdx.x * kTile + threadIdx.x * kPer;
  T s = 0;
#pragma unroll
  for (int q = 0; q < kPer; ++q)
    if (base + q < n) s += in[base + q];
  T tot;
  blockExclusiveScan<T>(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void tileScan(const T* __restrict__ in, T* __restrict__ out,
                                                     const T* __restrict__ offs, int n) {
  const long base = (long)blockIdx.x * kTile + threadIdx.x * kPer;
  T vals[kPer];
  T s = 0;
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    vals[q] = (base + q < n) ? in[base + q] : 0;
    s += vals[q];
  }
  T tot;
  T ex = blockExclusiveScan<T>(s, &tot) + (offs ? offs[blockIdx.x] : 0);
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    if (base + q < n) out[base + q] = ex;
    ex += vals[q];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kThreads - 1) out[n] = ex;
}

template <typename T>
void scanImpl(hipStream_t st, const T* in, T* out, int n, std::vector<T*>& scratch, int depth) {
  const int tiles = blocksFor(n, kTile);
  if (tiles == 1) {
    hipLaunchKernelGGL(tileScan<T>, dim3(1), dim3(kThreads), 0, st, in, out, (const T*)nullptr, n);
    return;
  }
  T* sums = scratch[2 * depth];
  T* sumsScan = scratch[2 * depth + 1];
  hipLaunchKernelGGL(tileSums<T>, dim3(tiles), dim3(kThreads), 0, st, in, sums, n);
  scanImpl<T>(st, sums, sumsScan, tiles, scratch, depth + 1);
  hipLaunchKernelGGL(tileScan<T>, dim3(tiles), dim3(kThreads), 0, st, in, out, (const T*)sumsScan, n);
}

template <typename T>
void scanTop(Scene& s, const T* in, T* out, int n) {
  if (n <= 0) {
    DV_HIP(hipMemsetAsync(out, 0, sizeof(T), s.stream));
    return;
  }
  // scratch: per level, tile sums and their scan (levels shrink by 1024x)
  std::vector<long> sizes;
  long m = n;
  while (m > kTile) {
    m = blocksFor(m, kTile);
    sizes.push_back(m);
    sizes.push_back(m + 1);
  }
  long total = 0;
  for (long z : sizes) total += z;
  s.scanTmp.reserve((size_t)total * sizeof(T) / sizeof(int) + 16);
  std::vector<T*> scratch;
  T* base = reinterpret_cast<T*>(s.scanTmp.p);
  for (long z : sizes) {
    scratch.push_back(base);
    base += z;
  }
  scanImpl<T>(s.stream, in, out, n, scratch, 0);
}

}  // namespace

void exclusiveScan(Scene& s, const int* in, int* out, int n) { scanTop<int>(s, in, out, n); }
void exclusiveScanU(Scene& s, const unsigned* in, unsigned* out, int n) {
  scanTop<unsigned>(s, in, out, n);
}

}  // namespace dv
