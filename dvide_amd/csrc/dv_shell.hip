// Fixed ANCF shell meshes on the contact path (System::importMesh,
// System.cu:1932-2060).  A shell whose DOFs carry no inverse mass does not
// move, so its contact points (System.cu:494-499) are fixed spheres of half
// the shell's thickness: they join the bodies after the caller's, in the
// reference's collision-geometry order (element by element, i then k), in
// collision family -2, and collision detection, assembly and the solvers
// treat them as bodies.  Two things remain shell-specific:
//
//  - each point's position, from its element's nodal coordinates through the
//    shape functions (CollisionDetector.cu:50-69), computed here on the device;
//  - the contact force on the shell DOFs, the shell columns of D^T gamma / h
//    (System.cu:904-961, 569-583): node by node, the forces of the points of
//    its elements weighted by the same shape functions.
#include "dv_geom.h"
#include "dv_internal.h"

#include <vector>

namespace dv {
namespace {

constexpr int kT = 256;

// position of contact point j, written as body nUser + j of p: node by node,
// the position, x-gradient and y-gradient terms of each component
__global__ __launch_bounds__(kT) void kShellPoints(const double* __restrict__ nodes, const int* __restrict__ conn,
                                                   const int* __restrict__ ptEl, const double* __restrict__ S,
                                                   int nPts, int nUser, double* __restrict__ p) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nPts) return;
  const int e = ptEl[j];
  double x[3] = {0.0, 0.0, 0.0};
  for (int q = 0; q < 4; ++q) {
    const double* nd = nodes + 9 * (size_t)conn[4 * e + q];
    const double* w = S + 12 * (size_t)j + 3 * q;
    for (int c = 0; c < 3; ++c) x[c] += (w[0] * nd[c] + w[1] * nd[3 + c]) + w[2] * nd[6 + c];
  }
  for (int c = 0; c < 3; ++c) p[3 * ((size_t)nUser + j) + c] = x[c];
}

// contact force on DOF t = 9 n + 3 s + c of node n (s: 0 position, 1
// x-gradient, 2 y-gradient): the sum of S_{q,s} f_j[c] over the node's
// (point j, corner q) incidences, in ascending order
__global__ __launch_bounds__(kT) void kShellForces(const int* __restrict__ incStart, const int* __restrict__ inc,
                                                   const double* __restrict__ S, const double* __restrict__ fc,
                                                   int nUser, int nNodes, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 9 * nNodes) return;
  const int n = t / 9, sl = (t % 9) / 3, c = t % 3;
  double acc = 0.0;
  for (int k = incStart[n]; k < incStart[n + 1]; ++k) {
    const int j = inc[k] >> 2, q = inc[k] & 3;
    acc += S[12 * (size_t)j + 3 * q + sl] * fc[3 * ((size_t)nUser + j) + c];
  }
  out[t] = acc;
}

template <typename T>
void put(DevBuf<T>& d, const std::vector<T>& h, hipStream_t st) {
  d.reserve(std::max<size_t>(h.size(), 1));
  if (!h.empty()) DV_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, st));
}

}  // namespace

void shellAppendPoints(Scene& s) {
  const int nEl = (int)s.hShellConn.size() / 4, w = s.shellPoints;
  s.hPtEl.clear();
  s.hPtS.clear();
  for (int e = 0; e < nEl; ++e)
    for (int i = 1; i < w - 1; ++i)
      for (int k = 1; k < w - 1; ++k) {
        double S[12];
        shellShape((double)i / (double)(w - 1), (double)k / (double)(w - 1), s.hShellGeom[3 * e],
                   s.hShellGeom[3 * e + 1], S);
        s.hPtEl.push_back(e);
        s.hPtS.insert(s.hPtS.end(), S, S + 12);
        for (int c = 0; c < 3; ++c) {
          s.hPos.push_back(0.0);  // placed on the device by shellPlacePoints
          s.hVel.push_back(0.0);
        }
        s.hGeom.push_back(0.5 * s.hShellGeom[3 * e + 2]);
        s.hGeom.push_back(0.0);
        s.hGeom.push_back(0.0);
        s.hMass.push_back(0.0);
        s.hFixed.push_back(1);
        s.hFamily.push_back(-2);
      }
}

void shellPlacePoints(Scene& s) {
  const hipStream_t st = s.stream;
  const int nPts = (int)s.hPtEl.size();
  put(s.shellNodes, s.hShellNodes, st);
  put(s.shellConn, s.hShellConn, st);
  put(s.ptEl, s.hPtEl, st);
  put(s.ptS, s.hPtS, st);
  // per node, its (point, corner) incidences in ascending point order
  std::vector<int> cnt(s.nNodes + 1, 0);
  for (int j = 0; j < nPts; ++j)
    for (int q = 0; q < 4; ++q) ++cnt[s.hShellConn[4 * s.hPtEl[j] + q] + 1];
  for (int n = 0; n < s.nNodes; ++n) cnt[n + 1] += cnt[n];
  std::vector<int> inc(cnt[s.nNodes]), fill(cnt.begin(), cnt.end() - 1);
  for (int j = 0; j < nPts; ++j)
    for (int q = 0; q < 4; ++q) inc[fill[s.hShellConn[4 * s.hPtEl[j] + q]]++] = (j << 2) | q;
  put(s.nodeIncStart, cnt, st);
  put(s.nodeInc, inc, st);
  s.shellF.reserve(std::max<size_t>(9 * (size_t)s.nNodes, 1));
  if (nPts)
    hipLaunchKernelGGL(kShellPoints, dim3((nPts + kT - 1) / kT), dim3(kT), 0, st, s.shellNodes.p, s.shellConn.p,
                       s.ptEl.p, s.ptS.p, nPts, s.nUser, s.p.p);
  DV_HIP(hipGetLastError());
}

void shellForces(Scene& s, double* out) {
  const int n = 9 * s.nNodes;
  if (n == 0) return;
  hipLaunchKernelGGL(kShellForces, dim3((n + kT - 1) / kT), dim3(kT), 0, s.stream, s.nodeIncStart.p, s.nodeInc.p,
                     s.ptS.p, s.fContact.p, s.nUser, s.nNodes, out);
  DV_HIP(hipGetLastError());
}

}  // namespace dv
