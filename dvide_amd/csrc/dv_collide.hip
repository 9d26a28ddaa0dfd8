// Collision detection for gfx950: a spatial-hash broadphase fused with the
// narrowphase, emitting contacts directly in canonical (idA, idB) order, plus
// the body-major incidence lists used for D^T products.
//
// Reference behaviour (CollisionDetector.cu):
//   * AABBs are geometry +- (radius or half extents + envelope)      (:24-79)
//   * candidate pairs exclude the same body and equal families != -1 (:131-132)
//   * the kept set is {pen >= -envelope} of the sphere-sphere and
//     sphere-box narrowphase; box-box never collides                 (:287-328)
//   * contacts are ordered by the 64-bit key (idA << 32) | idB after a
//     stable sort + unique of the candidates                           (:166-173, 621-622)
// The kept set does not depend on the bins (every kept pair's AABBs overlap
// with a margin of one envelope), so any complete broadphase yields the same
// contacts.  Here:
//   1. small spheres are hashed into cells of edge 2*(r_max + envelope), so a
//      kept partner lies in one of the 27 neighbouring cells; boxes and large
//      spheres form a short "big" list tested against every geometry;
//   2. one pass counts, per geometry, partners with a higher index (its row
//      of contacts) and a lower index (incidences where it is body B);
//   3. prefix sums give each geometry's contact row offset, so the emit pass
//      writes row i's contacts at their final canonical positions after a
//      thread-local sort of the row (no global key sort);
//   4. each body's incidences are listed in ascending contact order: first
//      the contacts where it is B, then its own row.
// The keep test and the contact record are evaluated with the shared
// dv_geom.h code (no fused multiply-add), bit-identical to the oracle.
#include "dv_geom.h"
#include "dv_internal.h"

#include <algorithm>
#include <cmath>

namespace dv {
namespace {

constexpr int kT = 256;

struct CdArgs {
  const double* p;
  const double* geom;
  const int* family;
  const int* smallIds;
  const int* bigIds;
  int nb, nsmall, nbig;
  double cs, envelope, rmaxSmall;
  unsigned mask;
  int mode;
  const unsigned* slotStart;
  const int* cellItems;
};

__device__ __forceinline__ unsigned cellHash(int cx, int cy, int cz, unsigned mask) {
  unsigned h = ((unsigned)cx * 73856093u) ^ ((unsigned)cy * 19349663u) ^ ((unsigned)cz * 83492791u);
  return h & mask;
}

__device__ __forceinline__ int cellCoord(double x, double cs) {
  double c = floor(x / cs);
  c = fmin(fmax(c, -1073741824.0), 1073741824.0);
  return (int)c;
}

__device__ __forceinline__ d3 ldPos(const double* p, int g) {
  return mk3(p[3 * g], p[3 * g + 1], p[3 * g + 2]);
}

// keep test of the pair {a, b}, evaluated in canonical order (min, max)
__device__ __forceinline__ bool keepPair(const CdArgs& a, int x, int y) {
  int lo = x < y ? x : y, hi = x < y ? y : x;
  int fl = a.family[lo], fh = a.family[hi];
  if (fl == fh && fl != -1) return false;
  d3 pA = ldPos(a.p, lo), pB = ldPos(a.p, hi);
  d3 gA = ldPos(a.geom, lo), gB = ldPos(a.geom, hi);
  return narrowphase(pA, gA, pB, gB, a.envelope, a.mode, (d3*)0, (double*)0, (d3*)0) != 0;
}

// Visit every candidate partner j != g of small geometry g: the occupants of
// the (deduplicated) hash slots of its 27 neighbour cells, then the big list.
// Cells have edge cs = 2 (rmax + envelope) >= r_g + r_j + envelope, the
// largest centre distance of a kept pair, so a kept partner lies in one of
// the 27 cells around g's cell.
template <class F>
__device__ __forceinline__ void forEachCandidate(const CdArgs& a, int g, F&& fn) {
  d3 pg = ldPos(a.p, g);
  int cx = cellCoord(pg.x, a.cs), cy = cellCoord(pg.y, a.cs), cz = cellCoord(pg.z, a.cs);
  unsigned slots[27];
#pragma unroll
  for (int q = 0; q < 27; ++q)
    slots[q] = cellHash(cx + q % 3 - 1, cy + (q / 3) % 3 - 1, cz + q / 9 - 1, a.mask);
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    bool dup = false;
#pragma unroll
    for (int w = 0; w < q; ++w) dup |= (slots[w] == slots[q]);
    if (dup) continue;
    unsigned s0 = a.slotStart[slots[q]], s1 = a.slotStart[slots[q] + 1];
    for (unsigned k = s0; k < s1; ++k) {
      int j = a.cellItems[k];
      if (j != g) fn(j);
    }
  }
  for (int q = 0; q < a.nbig; ++q) fn(a.bigIds[q]);
}

__global__ __launch_bounds__(kT) void fillU(unsigned* x, long n, unsigned v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

__global__ __launch_bounds__(kT) void hashSmall(CdArgs a, unsigned* slotOf, unsigned* slotCount) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nsmall) return;
  int g = a.smallIds[t];
  d3 pg = ldPos(a.p, g);
  unsigned s = cellHash(cellCoord(pg.x, a.cs), cellCoord(pg.y, a.cs), cellCoord(pg.z, a.cs), a.mask);
  slotOf[t] = s;
  atomicAdd(&slotCount[s], 1u);
}

__global__ __launch_bounds__(kT) void scatterSmall(CdArgs a, const unsigned* slotOf, unsigned* slotFill,
                                                   int* cellItems) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nsmall) return;
  unsigned s = slotOf[t];
  unsigned pos = a.slotStart[s] + atomicAdd(&slotFill[s], 1u);
  cellItems[pos] = a.smallIds[t];
}

// per small geometry: partners above (its contact row) and below
__global__ __launch_bounds__(kT) void countSmall(CdArgs a, int* cntUp, int* cntDown) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nsmall) return;
  int g = a.cellItems[t];  // spatial order: neighbouring threads share cells
  int up = 0, down = 0;
  forEachCandidate(a, g, [&](int j) {
    if (keepPair(a, g, j)) {
      if (j > g) ++up; else ++down;
    }
  });
  cntUp[g] = up;
  cntDown[g] = down;
}

__device__ int blockSum(int v) {
  __shared__ int red[kT / 64];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  int s = 0;
#pragma unroll
  for (int w = 0; w < kT / 64; ++w) s += red[w];
  __syncthreads();
  return s;
}

// Big geometries (boxes, large spheres) are paired with every geometry in
// chunks of kChunk partners: block (chunk, b) counts the kept partners of big
// geometry b in its chunk; a per-big scan over the chunks then gives each
// block its output offsets, so the emit pass writes the row in order.
constexpr int kChunk = 2048;

__global__ __launch_bounds__(kT) void countBigChunks(CdArgs a, int nchunks, int* chUp, int* chDown) {
  const int b = a.bigIds[blockIdx.y];
  const int j0 = blockIdx.x * kChunk, j1 = min(a.nb, j0 + kChunk);
  int up = 0, down = 0;
  for (int j = j0 + threadIdx.x; j < j1; j += kT) {
    if (j == b) continue;
    if (keepPair(a, b, j)) {
      if (j > b) ++up; else ++down;
    }
  }
  up = blockSum(up);
  down = blockSum(down);
  if (threadIdx.x == 0) {
    chUp[blockIdx.y * nchunks + blockIdx.x] = up;
    chDown[blockIdx.y * nchunks + blockIdx.x] = down;
  }
}

// per big geometry: exclusive scan of its chunk counts (in place) and totals
__global__ void scanBigChunks(CdArgs a, int nchunks, int* chUp, int* chDown, int* cntUp, int* cntDown) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.nbig) return;
  int su = 0, sd = 0;
  for (int k = 0; k < nchunks; ++k) {
    const int u = chUp[q * nchunks + k], d = chDown[q * nchunks + k];
    chUp[q * nchunks + k] = su;
    chDown[q * nchunks + k] = sd;
    su += u;
    sd += d;
  }
  cntUp[a.bigIds[q]] = su;
  cntDown[a.bigIds[q]] = sd;
}

__global__ __launch_bounds__(kT) void addCounts(const int* cntUp, const int* cntDown, int* tot, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) tot[i] = cntUp[i] + cntDown[i];
}

struct EmitOut {
  int* cA;
  int* cB;
  double* frame;  // 9 SoA, stride cap
  double* gap;
  double* loc;    // 3 SoA, stride cap
  int cap;
  int* inc;       // partner ids of lower-index contacts, converted later
  const int* cStart;
  const int* iStart;
  const int* cntUp;
  const int* cntDown;
};

__device__ __forceinline__ void writeRecord(const CdArgs& a, const EmitOut& o, int c, int A, int B) {
  d3 n, l;
  double w;
  narrowphase(ldPos(a.p, A), ldPos(a.geom, A), ldPos(a.p, B), ldPos(a.geom, B), a.envelope, a.mode,
              &n, &w, &l);
  d3 u, v;
  contactFrame(n, a.mode, &u, &v);
  const long S = o.cap;
  o.frame[0 * S + c] = n.x; o.frame[1 * S + c] = n.y; o.frame[2 * S + c] = n.z;
  o.frame[3 * S + c] = u.x; o.frame[4 * S + c] = u.y; o.frame[5 * S + c] = u.z;
  o.frame[6 * S + c] = v.x; o.frame[7 * S + c] = v.y; o.frame[8 * S + c] = v.z;
  o.gap[c] = w;
  o.loc[0 * S + c] = l.x; o.loc[1 * S + c] = l.y; o.loc[2 * S + c] = l.z;
  o.cA[c] = A;
  o.cB[c] = B;
}

__device__ void insertionSort(int* x, int n) {
  for (int i = 1; i < n; ++i) {
    int v = x[i], j = i - 1;
    while (j >= 0 && x[j] > v) {
      x[j + 1] = x[j];
      --j;
    }
    x[j + 1] = v;
  }
}

__global__ __launch_bounds__(kT) void emitSmall(CdArgs a, EmitOut o) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nsmall) return;
  int g = a.cellItems[t];
  const int c0 = o.cStart[g], i0 = o.iStart[g];
  int up = 0, down = 0;
  forEachCandidate(a, g, [&](int j) {
    if (keepPair(a, g, j)) {
      if (j > g) o.cB[c0 + up++] = j;
      else o.inc[i0 + down++] = j;
    }
  });
  insertionSort(o.cB + c0, up);
  insertionSort(o.inc + i0, down);
  for (int q = 0; q < up; ++q) writeRecord(a, o, c0 + q, g, o.cB[c0 + q]);
}

// big geometry rows: ordered stream compaction of each chunk's partners
__global__ __launch_bounds__(kT) void emitBigChunks(CdArgs a, EmitOut o, int nchunks, const int* chUp,
                                                    const int* chDown) {
  const int b = a.bigIds[blockIdx.y];
  const int j0 = blockIdx.x * kChunk, j1 = min(a.nb, j0 + kChunk);
  __shared__ int baseUp, baseDown;
  if (threadIdx.x == 0) {
    baseUp = o.cStart[b] + chUp[blockIdx.y * nchunks + blockIdx.x];
    baseDown = o.iStart[b] + chDown[blockIdx.y * nchunks + blockIdx.x];
  }
  __syncthreads();
  __shared__ int wUp[kT / 64], wDown[kT / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int jb = j0; jb < j1; jb += kT) {
    const int j = jb + threadIdx.x;
    const bool kept = (j < j1 && j != b) ? keepPair(a, b, j) : false;
    const bool isUp = kept && j > b, isDown = kept && j < b;
    const unsigned long long mUp = __ballot(isUp), mDown = __ballot(isDown);
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int rUp = __popcll(mUp & below), rDown = __popcll(mDown & below);
    if (lane == 0) { wUp[wave] = __popcll(mUp); wDown[wave] = __popcll(mDown); }
    __syncthreads();
    int offUp = baseUp, offDown = baseDown, totUp = 0, totDown = 0;
#pragma unroll
    for (int w = 0; w < kT / 64; ++w) {
      if (w < wave) { offUp += wUp[w]; offDown += wDown[w]; }
      totUp += wUp[w];
      totDown += wDown[w];
    }
    if (isUp) writeRecord(a, o, offUp + rUp, b, j);
    if (isDown) o.inc[offDown + rDown] = j;
    __syncthreads();
    if (threadIdx.x == 0) { baseUp += totUp; baseDown += totDown; }
    __syncthreads();
  }
}

// incidences of body g in ascending contact order: (c << 1) | side, side 1 =
// B (first the contacts of its lower-numbered partners' rows, then its own
// row); and each contact's slot in the B-side order, posB[c] = (iStart[g] -
// cStart[g]) + q for the q-th B incidence of g, so that the B incidences of
// every body are contiguous there
__global__ __launch_bounds__(kT) void buildIncidence(int nb, const int* cStart, const int* iStart,
                                                     const int* cntUp, const int* cntDown,
                                                     const int* cB, int* inc, int* posB, int* heavyFlag) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nb) return;
  const int i0 = iStart[g], nd = cntDown[g], nu = cntUp[g];
  const int b0 = i0 - cStart[g];
  for (int q = 0; q < nd; ++q) {
    int j = inc[i0 + q];
    int lo = cStart[j], hi = cStart[j] + cntUp[j];
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (cB[mid] < g) lo = mid + 1; else hi = mid;
    }
    inc[i0 + q] = (lo << 1) | 1;
    posB[lo] = b0 + q;
  }
  const int c0 = cStart[g];
  for (int q = 0; q < nu; ++q) inc[i0 + nd + q] = (c0 + q) << 1;
  heavyFlag[g] = nd + nu > kHeavyDegree ? 1 : 0;
}

// the bodies reduced by a whole workgroup, in ascending order (so that
// everything summed over them is summed in a fixed order), and their count
__global__ __launch_bounds__(kT) void compactHeavy(int nb, const int* heavyFlag, const int* pos, int* heavy,
                                                   int* counters) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g == 0) counters[2] = pos[nb];
  if (g < nb && heavyFlag[g]) heavy[pos[g]] = g;
}

__global__ void zeroInt(int* x, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0;
}

}  // namespace

void collideInit(Scene& s) {
  // classify: spheres up to 4x the median radius of the caller's spheres are
  // hashed, the rest (boxes, large spheres) are tested against every
  // geometry; the shells' contact points (bodies from nUser on, spheres of
  // half the shell's thickness) are always hashed
  std::vector<double> radii;
  for (int i = 0; i < s.nUser; ++i)
    if (s.hGeom[3 * i + 1] == 0) radii.push_back(s.hGeom[3 * i]);
  double rmed = 0;
  if (!radii.empty()) {
    std::vector<double> tmp = radii;
    std::nth_element(tmp.begin(), tmp.begin() + tmp.size() / 2, tmp.end());
    rmed = tmp[tmp.size() / 2];
  }
  std::vector<int> small, big;
  double rmax = 0;
  for (int i = 0; i < s.nb; ++i) {
    bool sphere = s.hGeom[3 * i + 1] == 0;
    if (sphere && (i >= s.nUser || s.hGeom[3 * i] <= 4.0 * rmed)) {
      small.push_back(i);
      rmax = std::max(rmax, s.hGeom[3 * i]);
    } else {
      big.push_back(i);
    }
  }
  s.nsmall = (int)small.size();
  s.nbig = (int)big.size();
  s.cellSize = rmax;  // scaled with the envelope at each step
  unsigned T = 1024;
  while (T < 2u * (unsigned)std::max(1, s.nsmall)) T <<= 1;
  s.hashSize = T;
  s.smallIds.reserve(std::max(1, s.nsmall));
  s.bigIds.reserve(std::max(1, s.nbig));
  if (s.nsmall)
    DV_HIP(hipMemcpyAsync(s.smallIds.p, small.data(), sizeof(int) * s.nsmall, hipMemcpyHostToDevice, s.stream));
  if (s.nbig)
    DV_HIP(hipMemcpyAsync(s.bigIds.p, big.data(), sizeof(int) * s.nbig, hipMemcpyHostToDevice, s.stream));
  s.slotOf.reserve(std::max(1, s.nsmall));
  s.slotCount.reserve(T);
  s.slotFill.reserve(T);
  s.slotStart.reserve(T + 1);
  s.cellItems.reserve(std::max(1, s.nsmall));
  s.cntUp.reserve(s.nb + 1);
  s.cntDown.reserve(s.nb + 1);
  s.cStart.reserve(s.nb + 1);
  s.iStart.reserve(s.nb + 1);
  s.counters.reserve(8);
  s.heavy.reserve(s.nb + 1);
  s.heavyPos.reserve(s.nb + 1);
  DV_HIP(hipStreamSynchronize(s.stream));
}

namespace {

CdArgs cdArgs(Scene& s) {
  CdArgs a;
  a.p = s.p.p;
  a.geom = s.geom.p;
  a.family = s.family.p;
  a.smallIds = s.smallIds.p;
  a.bigIds = s.bigIds.p;
  a.nb = s.nb;
  a.nsmall = s.nsmall;
  a.nbig = s.nbig;
  a.cs = 2.0 * (s.cellSize + s.envelope);
  a.rmaxSmall = s.cellSize;
  if (!(a.cs > 0)) a.cs = 1.0;
  a.envelope = s.envelope;
  a.mask = s.hashSize - 1;
  a.mode = s.geomMode;
  a.slotStart = s.slotStart.p;
  a.cellItems = s.cellItems.p;
  return a;
}

}  // namespace

void collideCount(Scene& s) {
  const hipStream_t st = s.stream;
  const CdArgs a = cdArgs(s);
  const unsigned T = s.hashSize;

  hipLaunchKernelGGL(zeroInt, dim3(1), dim3(64), 0, st, s.counters.p, 8);
  if (s.nsmall) {
    hipLaunchKernelGGL(fillU, dim3(blocksFor(T, kT)), dim3(kT), 0, st, s.slotCount.p, (long)T, 0u);
    hipLaunchKernelGGL(fillU, dim3(blocksFor(T, kT)), dim3(kT), 0, st, s.slotFill.p, (long)T, 0u);
    hipLaunchKernelGGL(hashSmall, dim3(blocksFor(s.nsmall, kT)), dim3(kT), 0, st, a, s.slotOf.p, s.slotCount.p);
    exclusiveScanU(s, s.slotCount.p, s.slotStart.p, (int)T);
    hipLaunchKernelGGL(scatterSmall, dim3(blocksFor(s.nsmall, kT)), dim3(kT), 0, st, a, s.slotOf.p,
                       s.slotFill.p, s.cellItems.p);
    hipLaunchKernelGGL(countSmall, dim3(blocksFor(s.nsmall, kT)), dim3(kT), 0, st, a, s.cntUp.p, s.cntDown.p);
  }
  const int nchunks = blocksFor(s.nb, kChunk);
  s.bigChunks.reserve((size_t)2 * std::max(1, s.nbig * nchunks));
  int* chUp = s.bigChunks.p;
  int* chDown = s.bigChunks.p + (size_t)std::max(1, s.nbig * nchunks);
  if (s.nbig) {
    hipLaunchKernelGGL(countBigChunks, dim3(nchunks, s.nbig), dim3(kT), 0, st, a, nchunks, chUp, chDown);
    hipLaunchKernelGGL(scanBigChunks, dim3(blocksFor(s.nbig, 64)), dim3(64), 0, st, a, nchunks, chUp, chDown,
                       s.cntUp.p, s.cntDown.p);
  }
  // contact rows and incidence offsets
  s.scanTmp.reserve(16);
  exclusiveScan(s, s.cntUp.p, s.cStart.p, s.nb);
  DevBuf<int>& tot = s.hostTotals;  // reuse as n-sized temp of up+down
  tot.reserve(s.nb + 1);
  hipLaunchKernelGGL(addCounts, dim3(blocksFor(s.nb, kT)), dim3(kT), 0, st, s.cntUp.p, s.cntDown.p, tot.p, s.nb);
  exclusiveScan(s, tot.p, s.iStart.p, s.nb);
  DV_HIP(hipMemcpyAsync(s.pinned, s.cStart.p + s.nb, sizeof(int), hipMemcpyDeviceToHost, st));
  DV_HIP(hipMemcpyAsync(s.pinned + 1, s.iStart.p + s.nb, sizeof(int), hipMemcpyDeviceToHost, st));
}

void collideEmit(Scene& s) {
  const hipStream_t st = s.stream;
  const CdArgs a = cdArgs(s);
  const int nchunks = blocksFor(s.nb, kChunk);
  int* chUp = s.bigChunks.p;
  int* chDown = s.bigChunks.p + (size_t)std::max(1, s.nbig * nchunks);
  DV_HIP(hipStreamSynchronize(st));  // the counts of collideCount
  const int nc = s.pinned[0];
  const int ninc = s.pinned[1];
  if (nc < 0 || ninc < 0 || nc > kMaxContacts)
    throw DvError{DV_ERR_UNSUPPORTED, "more than 2^28 contacts in one scene"};
  if (ninc != 2 * nc) throw DvError{DV_ERR_STATE, "collision detection: asymmetric pair counts"};
  s.nc = nc;
  if (nc > s.ncCap || s.ncCap == 0) {
    int cap = std::max(1024, nc + nc / 4);
    s.cA.release(); s.cB.release(); s.frame.release(); s.gap.release(); s.loc.release();
    s.frame32.release(); s.inc.release(); s.posB.release();
    s.cA.reserve(cap); s.cB.reserve(cap);
    s.frame.reserve((size_t)9 * cap);
    s.gap.reserve(cap);
    s.loc.reserve((size_t)3 * cap);
    s.inc.reserve((size_t)2 * cap);
    s.posB.reserve(cap);
    s.ncCap = cap;
  }
  EmitOut o;
  o.cA = s.cA.p; o.cB = s.cB.p; o.frame = s.frame.p; o.gap = s.gap.p; o.loc = s.loc.p;
  o.cap = s.ncCap;
  o.inc = s.inc.p;
  o.cStart = s.cStart.p; o.iStart = s.iStart.p; o.cntUp = s.cntUp.p; o.cntDown = s.cntDown.p;
  if (s.nsmall) hipLaunchKernelGGL(emitSmall, dim3(blocksFor(s.nsmall, kT)), dim3(kT), 0, st, a, o);
  if (s.nbig)
    hipLaunchKernelGGL(emitBigChunks, dim3(nchunks, s.nbig), dim3(kT), 0, st, a, o, nchunks, chUp, chDown);
  int* heavyFlag = s.hostTotals.p;  // free again after the offsets' scans
  hipLaunchKernelGGL(buildIncidence, dim3(blocksFor(s.nb, kT)), dim3(kT), 0, st, s.nb, s.cStart.p,
                     s.iStart.p, s.cntUp.p, s.cntDown.p, s.cB.p, s.inc.p, s.posB.p, heavyFlag);
  exclusiveScan(s, heavyFlag, s.heavyPos.p, s.nb);
  hipLaunchKernelGGL(compactHeavy, dim3(blocksFor(s.nb, kT)), dim3(kT), 0, st, s.nb, (const int*)heavyFlag,
                     (const int*)s.heavyPos.p, s.heavy.p, s.counters.p);
  DV_HIP(hipGetLastError());
}

void collide(Scene& s) {
  collideCount(s);
  collideEmit(s);
}

}  // namespace dv
