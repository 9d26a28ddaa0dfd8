// APGD (APGD::solve, APGD.cu:164-361) as a device-resident state machine
// driven by a captured hipGraph: no host round trip per reduction.
//
// One solver pass (one Schur product plus the vector algebra around it) is
// three kernels:
//   P1 (contacts): first the decision of the previous pass -- fixed-order
//       reduction of its per-workgroup partials, the Lipschitz test
//       (APGD.cu:236-264) and either a backtracking step or the acceptance
//       logic (APGD.cu:266-332) -- made redundantly by every workgroup from
//       the same inputs; then y and N y from the last two iterates,
//       g = N y + r, gamma_new = Pi(y - t g), the dot products of the
//       objective and restart tests, and J_c = F_c^T gamma_new (the
//       world-frame impulse of contact c on body A);
//   P2 (bodies):   tmp_b = M_b^-1 (sum_A J_c - sum_B J_c + bilateral rows);
//       bodies with many contacts are reduced by a whole workgroup;
//   P3 (contacts): (N gamma_new)_c = F_c (tmp_A - tmp_B), gamma_new . N gamma_new
//       and the residual's infinity norm (getResidual, APGD.cu:90-102).
// State and partials are double-buffered by pass parity so that the
// redundant decisions never race with their publication.  A one-wavefront
// kernel decides the last pass of each chunk.
// Kernel boundaries are the grid-wide barriers: on MI355X a kernel boundary
// inside a graph costs ~1.6 us, less than a grid barrier of a persistent
// cooperative kernel (3.7 us at 128 workgroups, ~12 us at 512).  The host
// launches the graph of a chunk of passes, reads one flag, and repeats; passes
// after convergence exit immediately.
//
// The same arithmetic as the reference, less traffic:
//   * N y is not recomputed: with y_{k+1} = (1+beta) g_{k+1} - beta g_k,
//     N y_{k+1} = (1+beta) N g_{k+1} - beta N g_k (N g on restart);
//   * N gamma_new (APGD.cu:237) and the product in getResidual (APGD.cu:276)
//     are the same vector, so each pass performs one product;
//   * a contact's tangent frame is rebuilt from its normal (System.cu:753-762)
//     instead of being read.
// Layout: row vectors are stored component-major (all n rows, then all u
// rows, then all v rows, then the bilateral rows), so every wavefront load is
// one contiguous 512-byte (FP64) run; J and tmp, which are gathered through
// the incidence lists, are padded to four components for one aligned load
// per gather.  The reference's flat layout (bilateral rows first, then n,u,v
// per contact) is used only at the boundary (gamma in/out, r in).
//
// TM is the storage/multiply type of D (frames), M^-1, r and the products
// (J, tmp, N gamma): double for the FP64 solver, float for the mixed solver.
// The iterates gamma, y and the step gamma_new = Pi(y - t g) are FP64 in
// both: with binary32 iterates the small projected-gradient steps round away
// and APGD stalls near a residual of 1e-2, while binary32 products with FP64
// iterates converge like the FP64 solver.  Dot products accumulate in FP64.
// The mixed solver evaluates the Lipschitz test in the cancellation-free form
// D'N D >= L |D|^2 (D = gamma_new - y), equal to obj1 >= obj2 for g = N y + r
// and symmetric N, because its binary32 N gamma would swamp the small
// difference obj2 - obj1.
#include "dv_geom.h"
#include "dv_internal.h"

#include <algorithm>
#include <cstdio>

namespace dv {

struct ApgdState {
  double L, t, theta, beta, best, objective;
  double sum[2 * kSlots];
  int it, bt, backtracks, restart;
  int cur, prev, nxt;
  int pending;  // partials of a pass are waiting for their decision
  int hatSrc, done, converged, passes;
};

struct SolveDims {
  int nc, nbil, nb, maxIt;
  double mu, tol, gdiff;
  int warm, geomMode;
};

namespace {

constexpr int kTC = 1024;  // contact kernels: 16 wavefronts per workgroup
constexpr int kTB = 256;   // body kernel
constexpr int kTD = 64;    // decide kernels: one wavefront
constexpr int kBacktrackCap = 1100;  // L doubles past 1e308 long before this
constexpr int kPassesPerChunk = 16;  // even: a chunk's final state lands in st[0]
static_assert(kPassesPerChunk % 2 == 0, "pass parity must return to 0 at a chunk end");
constexpr int kMaxGridC = 256;  // reduceAll handles up to 4*64 partials

template <typename TM>
struct V4 {
  TM x, y, z, w;
};

template <typename TM>
struct Args {
  const SolveDims* dims;
  ApgdState* st;
  const TM* frame;      // SoA stride cap: n (FP64) or n, u (FP32)
  long cap;             // contact capacity: row-vector stride
  const int* cA;
  const int* cB;
  const int* iStart;
  const int* inc;
  const int* bilIncStart;
  const int* bilInc;
  const int* bilA;
  const int* bilB;
  const int* heavy;
  const int* counters;  // [2] = number of heavy bodies
  const TM* minv;
  const double* rFlat;  // assembled r, reference layout
  TM* r;                // r, component-major
  double* gammaIO;      // flat gamma: previous values in, result out
  double* ring[4];      // iterates (component-major), slots named by ApgdState
  TM* rn[4];            // N of the iterate in the same ring slot
  V4<TM>* J;
  V4<TM>* tmp;
  double* gHat;         // gammaHat before any iterate is accepted (ones)
  TM* delta;            // mixed solver: gamma_new - y
  double* part1[2];     // kSlots per workgroup, by pass parity
  double* part3[2];
};

// row offsets: contact c component k -> k*cap + c; bilateral b -> 3*cap + b
#define ROW(a, k, c) ((long)(k) * (a).cap + (c))
#define BROW(a, b) (3L * (a).cap + (b))

// ---------------------------------------------------------------------------
// reductions (slot kSlots-1 is a max, the others sums)

__device__ __forceinline__ double waveSum(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}
__device__ __forceinline__ double waveMax(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = fmax(v, __shfl_xor(v, d, 64));
  return v;
}

template <int NT, int NS>
__device__ void blockReduce(double (&acc)[NS], double* out) {
  constexpr int W = NT / 64;
  __shared__ double sh[W][NS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    double v = (s % kSlots == kSlots - 1) ? waveMax(acc[s]) : waveSum(acc[s]);
    if (lane == 0) sh[wave][s] = v;
  }
  __syncthreads();
  if (threadIdx.x < NS) {
    const int s = threadIdx.x;
    double v = sh[0][s];
    for (int w = 1; w < W; ++w) v = (s % kSlots == kSlots - 1) ? fmax(v, sh[w][s]) : v + sh[w][s];
    out[s] = v;
  }
  __syncthreads();
}

template <int NT>
__device__ void writePartials(double (&acc)[kSlots], double* part) {
  __shared__ double tot[kSlots];
  blockReduce<NT, kSlots>(acc, tot);
  if (threadIdx.x < kSlots) part[(long)blockIdx.x * kSlots + threadIdx.x] = tot[threadIdx.x];
}

// fixed-order totals of the partials of `ng` <= 4*64 workgroups from P1 and
// P3, by one wavefront: out[0..7] from part1, out[8..15] from part3 (valid
// in every lane).  All loads are issued before any is consumed.
__device__ void reduceAll(const double* part1, const double* part3, int ng, double* out) {
  const int lane = threadIdx.x;
  double x[4][2 * kSlots];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int w = lane + 64 * k;
    const int ws = w < ng ? w : 0;
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
      x[k][s] = part1[(long)ws * kSlots + s];
      x[k][kSlots + s] = part3[(long)ws * kSlots + s];
    }
  }
#pragma unroll
  for (int s = 0; s < 2 * kSlots; ++s) {
    const bool isMax = (s % kSlots == kSlots - 1);
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double xv = (lane + 64 * k) < ng ? x[k][s] : 0.0;
      v = isMax ? fmax(v, xv) : v + xv;
    }
    out[s] = isMax ? waveMax(v) : waveSum(v);
  }
}

// ---------------------------------------------------------------------------
// contact frames and row helpers

__device__ __forceinline__ void ldFrame(const Args<double>& a, int c, int mode, double (&F)[9]) {
  const long S = a.cap;
  const d3 n = mk3(a.frame[c], a.frame[S + c], a.frame[2 * S + c]);
  d3 u, v;
  contactFrame(n, mode, &u, &v);  // bit-identical to the assembled frame
  F[0] = n.x; F[1] = n.y; F[2] = n.z;
  F[3] = u.x; F[4] = u.y; F[5] = u.z;
  F[6] = v.x; F[7] = v.y; F[8] = v.z;
}

__device__ __forceinline__ void ldFrame(const Args<float>& a, int c, int, float (&F)[9]) {
  const long S = a.cap;
#pragma unroll
  for (int k = 0; k < 6; ++k) F[k] = a.frame[k * S + c];
  F[6] = F[1] * F[5] - F[2] * F[4];  // v = n x u
  F[7] = F[2] * F[3] - F[0] * F[5];
  F[8] = F[0] * F[4] - F[1] * F[3];
}

template <typename TM>
__device__ __forceinline__ void storeJ(const Args<TM>& a, int c, const TM (&F)[9], TM x0, TM x1, TM x2) {
  V4<TM> j;
  j.x = F[0] * x0 + F[3] * x1 + F[6] * x2;
  j.y = F[1] * x0 + F[4] * x1 + F[7] * x2;
  j.z = F[2] * x0 + F[5] * x1 + F[8] * x2;
  j.w = 0;
  a.J[c] = j;
}

template <typename TM>
__device__ __forceinline__ TM tmpDof(const Args<TM>& a, int dof) {
  const V4<TM> t = a.tmp[dof / 3];
  const int k = dof % 3;
  return k == 0 ? t.x : (k == 1 ? t.y : t.z);
}

template <typename TM>
__device__ __forceinline__ TM bilRow(const Args<TM>& a, int q) {
  TM s = tmpDof(a, a.bilA[q]);
  if (a.bilB[q] >= 0) s -= tmpDof(a, a.bilB[q]);
  return s;
}

template <typename TM>
__device__ __forceinline__ void contactRows(const Args<TM>& a, int c, const TM (&F)[9], TM& n0, TM& n1, TM& n2) {
  const V4<TM> tA = a.tmp[a.cA[c]], tB = a.tmp[a.cB[c]];
  const TM d0 = tA.x - tB.x, d1 = tA.y - tB.y, d2 = tA.z - tB.z;
  n0 = F[0] * d0 + F[1] * d1 + F[2] * d2;
  n1 = F[3] * d0 + F[4] * d1 + F[5] * d2;
  n2 = F[6] * d0 + F[7] * d1 + F[8] * d2;
}

__device__ __forceinline__ double residualContact(double g0, double g1, double g2, double w0, double w1,
                                                  double w2, double mu, double gdiff) {
  double z0 = 1.0 * g0 + (-gdiff) * w0, z1 = 1.0 * g1 + (-gdiff) * w1, z2 = 1.0 * g2 + (-gdiff) * w2;
  projectCone<double>(z0, z1, z2, mu);
  const double ig = 1.0 / gdiff;
  const double e0 = ig * g0 + (-ig) * z0, e1 = ig * g1 + (-ig) * z1, e2 = ig * g2 + (-ig) * z2;
  return fmax(fabs(e0), fmax(fabs(e1), fabs(e2)));
}

__device__ __forceinline__ void zero(double (&acc)[kSlots]) {
#pragma unroll
  for (int s = 0; s < kSlots; ++s) acc[s] = 0.0;
}

#define GRID_LOOP(q, n, NT) for (long q = (long)blockIdx.x * NT + threadIdx.x; q < (n); q += (long)gridDim.x * NT)

// ---------------------------------------------------------------------------
// setup (APGD.cu:195-223)

// r to component-major; gamma_0, gammaHat_0 = 1, d = gamma_0 - gammaHat_0 (in
// ring[1]); J = F^T d; |d|^2
template <typename TM>
__global__ __launch_bounds__(kTC) void kSetup1(Args<TM> a) {
  const SolveDims D = *a.dims;
  const long items = (long)D.nc + D.nbil, n3 = 3L * D.nc;
  double acc[kSlots];
  zero(acc);
  GRID_LOOP(q, items, kTC) {
    const bool bil = q >= D.nc;
    const int nr = bil ? 1 : 3;
    double x[3] = {0, 0, 0};
    for (int rr = 0; rr < nr; ++rr) {
      const long f = bil ? (q - D.nc) : (D.nbil + 3L * q + rr);  // flat (reference) row
      const long o = bil ? BROW(a, q - D.nc) : ROW(a, rr, q);
      a.r[o] = (TM)a.rFlat[f];
      // flat entries 0..3nc-1 start as (1,0,0) triples, without the
      // bilateral offset (APGD.cu:116-122); the rest keep their old values
      const double gv = (!D.warm && f < n3) ? ((f % 3 == 0) ? 1.0 : 0.0) : a.gammaIO[f];
      a.ring[0][o] = gv;
      a.gHat[o] = 1.0;
      x[rr] = 1.0 * gv + (-1.0) * 1.0;
      a.ring[1][o] = x[rr];
      acc[0] += x[rr] * x[rr];
    }
    if (!bil) {
      TM F[9];
      ldFrame(a, (int)q, D.geomMode, F);
      storeJ(a, (int)q, F, (TM)x[0], (TM)x[1], (TM)x[2]);
    }
  }
  writePartials<kTC>(acc, a.part1[0]);
}

// P2: tmp = M^-1 D^T x for non-fixed bodies; X holds the bilateral rows of x.
// which: 0 = ring[1] (d), 1 = ring[0] (gamma_0), 2 = ring[nxt] (gamma_new of
// the pass whose state is st[sIdx]), 3 = gHat
template <typename TM>
__global__ __launch_bounds__(kTB) void kBody(Args<TM> a, int which, int sIdx) {
  if (which == 2 && a.st[sIdx].done) return;
  const SolveDims D = *a.dims;
  const double* X = which == 0 ? a.ring[1] : which == 1 ? a.ring[0] : which == 2 ? a.ring[a.st[sIdx].nxt] : a.gHat;
  const long bil0 = 3L * a.cap;
  GRID_LOOP(g, D.nb, kTB) {
    const TM mi = a.minv[g];
    const int k0 = a.iStart[g], k1 = a.iStart[g + 1];
    V4<TM> out;
    out.x = 0; out.y = 0; out.z = 0; out.w = 0;
    if (mi != (TM)0) {
      if (k1 - k0 > kHeavyDegree) continue;  // reduced below by a workgroup
      TM s0 = 0, s1 = 0, s2 = 0;
      for (int k = a.bilIncStart[g]; k < a.bilIncStart[g + 1]; ++k) {
        const int e = a.bilInc[k];
        const TM t = (TM)((e & 1) ? -X[bil0 + (e >> 3)] : X[bil0 + (e >> 3)]);
        const int comp = (e >> 1) & 3;
        if (comp == 0) s0 += t; else if (comp == 1) s1 += t; else s2 += t;
      }
      // incidences in batches of 8: all index loads, then all J loads, then
      // the sums (in incidence order), so the gathers overlap
      for (int kb = k0; kb < k1; kb += 8) {
        int e[8];
        V4<TM> j[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) e[u] = (kb + u < k1) ? a.inc[kb + u] : -1;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          if (e[u] >= 0) j[u] = a.J[e[u] >> 1];
          else { j[u].x = 0; j[u].y = 0; j[u].z = 0; j[u].w = 0; }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          if (e[u] < 0) continue;
          if (e[u] & 1) { s0 -= j[u].x; s1 -= j[u].y; s2 -= j[u].z; }
          else { s0 += j[u].x; s1 += j[u].y; s2 += j[u].z; }
        }
      }
      out.x = mi * s0; out.y = mi * s1; out.z = mi * s2;
    }
    a.tmp[g] = out;
  }
  const int nHeavy = a.counters[2];
  for (int hq = blockIdx.x; hq < nHeavy; hq += gridDim.x) {
    const int g = a.heavy[hq];
    const TM mi = a.minv[g];
    if (mi == (TM)0) continue;
    const int k0 = a.iStart[g], k1 = a.iStart[g + 1];
    double acc[kSlots];
    zero(acc);
    for (int k = k0 + threadIdx.x; k < k1; k += kTB) {
      const int e = a.inc[k];
      const V4<TM> j = a.J[e >> 1];
      const double sg = (e & 1) ? -1.0 : 1.0;
      acc[0] += sg * (double)j.x;
      acc[1] += sg * (double)j.y;
      acc[2] += sg * (double)j.z;
    }
    if (threadIdx.x == 0) {
      for (int k = a.bilIncStart[g]; k < a.bilIncStart[g + 1]; ++k) {
        const int e = a.bilInc[k];
        acc[(e >> 1) & 3] += (e & 1) ? -X[bil0 + (e >> 3)] : X[bil0 + (e >> 3)];
      }
    }
    __shared__ double tot[kSlots];
    blockReduce<kTB, kSlots>(acc, tot);
    if (threadIdx.x == 0) {
      V4<TM> out;
      out.x = mi * (TM)tot[0];
      out.y = mi * (TM)tot[1];
      out.z = mi * (TM)tot[2];
      out.w = 0;
      a.tmp[g] = out;
    }
  }
}

// |N d|^2 into part3[1]; J = F^T gamma_0
template <typename TM>
__global__ __launch_bounds__(kTC) void kSetup3(Args<TM> a) {
  const SolveDims D = *a.dims;
  const long items = (long)D.nc + D.nbil;
  double acc[kSlots];
  zero(acc);
  GRID_LOOP(q, items, kTC) {
    if (q >= D.nc) {
      const double v = (double)bilRow(a, (int)(q - D.nc));
      acc[1] += v * v;
    } else {
      const int c = (int)q;
      TM F[9];
      ldFrame(a, c, D.geomMode, F);
      TM n0, n1, n2;
      contactRows(a, c, F, n0, n1, n2);
      acc[1] += (double)n0 * (double)n0 + (double)n1 * (double)n1 + (double)n2 * (double)n2;
      storeJ(a, c, F, (TM)a.ring[0][ROW(a, 0, c)], (TM)a.ring[0][ROW(a, 1, c)], (TM)a.ring[0][ROW(a, 2, c)]);
    }
  }
  writePartials<kTC>(acc, a.part3[0]);
}

// N gamma_0 into rn[0]
template <typename TM>
__global__ __launch_bounds__(kTC) void kSetup5(Args<TM> a) {
  const SolveDims D = *a.dims;
  const long items = (long)D.nc + D.nbil;
  GRID_LOOP(q, items, kTC) {
    if (q >= D.nc) {
      a.rn[0][BROW(a, q - D.nc)] = bilRow(a, (int)(q - D.nc));
    } else {
      const int c = (int)q;
      TM F[9];
      ldFrame(a, c, D.geomMode, F);
      TM n0, n1, n2;
      contactRows(a, c, F, n0, n1, n2);
      a.rn[0][ROW(a, 0, c)] = n0; a.rn[0][ROW(a, 1, c)] = n1; a.rn[0][ROW(a, 2, c)] = n2;
    }
  }
}

// L_0 = |N d| / |d|, t_0 = 1/L_0 and the initial state (APGD.cu:208-223)
template <typename TM>
__global__ __launch_bounds__(kTD) void kSetupDecide(Args<TM> a, int ng) {
  double tot[2 * kSlots];
  reduceAll(a.part1[0], a.part3[0], ng, tot);
  if (threadIdx.x == 0) {
    ApgdState S;
    double L = sqrt(tot[0]);
    L = sqrt(tot[kSlots + 1]) / L;
    S.L = L;
    S.t = 1.0 / L;
    S.theta = 1.0;
    S.beta = 0.0;
    S.best = 10e30;
    S.objective = 0;
    for (int k = 0; k < 2 * kSlots; ++k) S.sum[k] = 0;
    S.it = 0; S.bt = 0; S.backtracks = 0; S.restart = 1;
    S.cur = 0; S.prev = 0; S.nxt = 1; S.pending = 0; S.hatSrc = -1;  // -1: gHat (ones)
    S.done = a.dims->maxIt <= 0 ? 1 : 0;
    S.converged = 0;
    S.passes = 0;
    a.st[0] = S;
  }
}

// ---------------------------------------------------------------------------
// the decision of a pass: Lipschitz test, backtracking or acceptance
// (APGD.cu:236-332), from the pass's fixed-order partial sums tot[0..15]

template <bool kMixed>
__device__ void decideStep(ApgdState& S, const double* tot, const SolveDims& D) {
  for (int k = 0; k < 2 * kSlots; ++k) S.sum[k] = tot[k];
  const double* s1 = tot;           // P1 slots
  const double* s3 = tot + kSlots;  // P3 slots
  S.passes++;
  const double L = S.L;
  bool backtrack;
  if (kMixed) {
    backtrack = (s3[6] - s1[1]) >= L * s1[4];
  } else {
    const double obj1 = 0.5 * s3[6] + s1[0];
    const double nrm = sqrt(s1[4]);
    double obj2 = 0.5 * s1[1] + s1[2];
    obj2 += s1[3] + 0.5 * L * (nrm * nrm);
    backtrack = obj1 >= obj2;
  }
  if (backtrack && S.bt < kBacktrackCap) {
    S.bt++;
    S.backtracks++;
    S.L = 2.0 * L;
    S.t = 1.0 / S.L;
    return;
  }
  const double theta = S.theta;
  double thetaNew = (-(theta * theta) + theta * sqrt(theta * theta + 4.0)) / 2.0;
  S.beta = theta * (1.0 - theta) / (theta * theta + thetaNew);
  const double res = s3[7];
  if (res < S.best) {  // gammaHat := gamma_new, kept in its ring slot
    S.best = res;
    S.hatSrc = S.nxt;
  }
  if (S.best < D.tol) {
    S.done = 1;
    S.converged = 1;
    return;
  }
  if (s1[5] > 0) {
    S.restart = 1;
    thetaNew = 1.0;
  } else {
    S.restart = 0;
  }
  S.L = 0.9 * L;
  S.t = 1.0 / S.L;
  S.theta = thetaNew;
  S.prev = S.cur;
  S.cur = S.nxt;
  // next slot: the one of the four not holding gamma_k, gamma_{k-1} or gammaHat
  for (int k = 0; k < 4; ++k)
    if (k != S.cur && k != S.prev && k != S.hatSrc) { S.nxt = k; break; }
  S.it++;
  S.bt = 0;
  if (S.it >= D.maxIt) S.done = 1;
}

// ---------------------------------------------------------------------------
// one pass: P1 (parity par) reads the state st[par] and, if a pass is
// pending, the partials part1/part3[par ^ 1] of that pass; wavefront 0 of
// every workgroup applies the decision to its copy (same inputs, same fixed
// reduction order, same result), block 0 publishes the new state in
// st[par ^ 1] for P2, P3 and the next pass, and P1 writes part1[par].

template <typename TM>
__device__ __forceinline__ ApgdState passState(const Args<TM>& a, int par, const SolveDims& D) {
  constexpr bool kMixed = sizeof(TM) < sizeof(double);
  __shared__ ApgdState sh;
  if (threadIdx.x < 64) {
    ApgdState S = a.st[par];
    if (S.pending && !S.done) {
      double tot[2 * kSlots];
      reduceAll(a.part1[par ^ 1], a.part3[par ^ 1], gridDim.x, tot);
      decideStep<kMixed>(S, tot, D);
    }
    S.pending = S.done ? 0 : 1;
    if (threadIdx.x == 0) {
      sh = S;
      if (blockIdx.x == 0) a.st[par ^ 1] = S;
    }
  }
  __syncthreads();
  return sh;
}

template <typename TM>
__global__ __launch_bounds__(kTC) void kP1(Args<TM> a, int par) {
  constexpr bool kMixed = sizeof(TM) < sizeof(double);
  const SolveDims D = *a.dims;
  const ApgdState S = passState(a, par, D);
  if (S.done) return;
  const long items = (long)D.nc + D.nbil;
  const double* Gk = a.ring[S.cur];
  const double* Gkm1 = a.ring[S.prev];
  const TM* Nk = a.rn[S.cur];
  const TM* Nkm1 = a.rn[S.prev];
  double* Gn = a.ring[S.nxt];
  const double bp = 1.0 + S.beta, bm = -S.beta;
  const double t = S.t;
  double acc[kSlots];
  zero(acc);
  GRID_LOOP(q, items, kTC) {
    const bool bil = q >= D.nc;
    const int nr = bil ? 1 : 3;
    double y[3] = {0, 0, 0}, ny[3] = {0, 0, 0}, gk[3] = {0, 0, 0}, gg[3] = {0, 0, 0}, gn[3] = {0, 0, 0},
           rv[3] = {0, 0, 0};
    long o[3] = {0, 0, 0};
#pragma unroll
    for (int rr = 0; rr < 3; ++rr) {
      if (rr < nr) {
        o[rr] = bil ? BROW(a, q - D.nc) : ROW(a, rr, q);
        const long f = o[rr];
        gk[rr] = Gk[f];
        rv[rr] = (double)a.r[f];
        if (S.restart) {
          y[rr] = gk[rr];
          ny[rr] = (double)Nk[f];
        } else {
          y[rr] = bp * gk[rr] + bm * Gkm1[f];
          ny[rr] = bp * (double)Nk[f] + bm * (double)Nkm1[f];
        }
        gg[rr] = 1.0 * ny[rr] + 1.0 * rv[rr];
        gn[rr] = 1.0 * y[rr] + (-t) * gg[rr];
      }
    }
    if (!bil) projectCone<double>(gn[0], gn[1], gn[2], D.mu);
#pragma unroll
    for (int rr = 0; rr < 3; ++rr) {
      if (rr < nr) {
        const long f = o[rr];
        Gn[f] = gn[rr];
        const double dy = gn[rr] - y[rr];
        if (kMixed) {
          a.delta[f] = (TM)dy;
          acc[1] += dy * ny[rr];
        } else {
          acc[0] += gn[rr] * rv[rr];
          acc[1] += y[rr] * ny[rr];
          acc[2] += y[rr] * rv[rr];
          acc[3] += gg[rr] * dy;
        }
        acc[4] += dy * dy;
        acc[5] += gg[rr] * (gn[rr] - gk[rr]);
      }
    }
    if (!bil) {
      TM F[9];
      ldFrame(a, (int)q, D.geomMode, F);
      storeJ(a, (int)q, F, (TM)gn[0], (TM)gn[1], (TM)gn[2]);
    }
  }
  writePartials<kTC>(acc, a.part1[par]);
}

template <typename TM>
__global__ __launch_bounds__(kTC) void kP3(Args<TM> a, int par) {
  constexpr bool kMixed = sizeof(TM) < sizeof(double);
  const ApgdState S = a.st[par ^ 1];
  if (S.done) return;
  const SolveDims D = *a.dims;
  const long items = (long)D.nc + D.nbil;
  const double* Gn = a.ring[S.nxt];
  TM* Nn = a.rn[S.nxt];
  double acc[kSlots];
  zero(acc);
  GRID_LOOP(q, items, kTC) {
    if (q >= D.nc) {
      const long f = BROW(a, q - D.nc);
      const TM nv = bilRow(a, (int)(q - D.nc));
      Nn[f] = nv;
      const double gv = Gn[f];
      acc[6] += (kMixed ? (double)a.delta[f] : gv) * (double)nv;
      const double w = (double)nv + (double)a.r[f];
      double e;
      if (D.nc) {
        const double z = 1.0 * gv + (-D.gdiff) * w;
        const double ig = 1.0 / D.gdiff;
        e = fabs(ig * gv + (-ig) * z);
      } else {
        e = fabs(w);
      }
      acc[7] = fmax(acc[7], e);
    } else {
      const int c = (int)q;
      TM F[9];
      ldFrame(a, c, D.geomMode, F);
      TM n0, n1, n2;
      contactRows(a, c, F, n0, n1, n2);
      const long f0 = ROW(a, 0, c), f1 = ROW(a, 1, c), f2 = ROW(a, 2, c);
      Nn[f0] = n0; Nn[f1] = n1; Nn[f2] = n2;
      const double g0 = Gn[f0], g1 = Gn[f1], g2 = Gn[f2];
      if (kMixed)
        acc[6] += (double)a.delta[f0] * (double)n0 + (double)a.delta[f1] * (double)n1 +
                  (double)a.delta[f2] * (double)n2;
      else
        acc[6] += g0 * (double)n0 + g1 * (double)n1 + g2 * (double)n2;
      const double w0 = (double)n0 + (double)a.r[f0];
      const double w1 = (double)n1 + (double)a.r[f1];
      const double w2 = (double)n2 + (double)a.r[f2];
      acc[7] = fmax(acc[7], residualContact(g0, g1, g2, w0, w1, w2, D.mu, D.gdiff));
    }
  }
  writePartials<kTC>(acc, a.part3[par]);
}

// Chunk-end decision: the state of the last pass of a chunk sits in
// st[par ^ 1] and its partials in part1/part3[par]; decide it, so that the
// host sees the final state and the next chunk's first P1 has nothing pending.
template <typename TM>
__global__ __launch_bounds__(kTD) void kDecide(Args<TM> a, int par, int ng) {
  constexpr bool kMixed = sizeof(TM) < sizeof(double);
  ApgdState S = a.st[par ^ 1];
  if (!S.pending) return;
  if (!S.done) {
    double tot[2 * kSlots];
    reduceAll(a.part1[par], a.part3[par], ng, tot);
    const SolveDims D = *a.dims;
    decideStep<kMixed>(S, tot, D);
  }
  S.pending = 0;
  if (threadIdx.x == 0) a.st[par ^ 1] = S;
}

// ---------------------------------------------------------------------------
// gamma := gammaHat (APGD.cu:358) and objectiveCCP (APGD.cu:353-354)

template <typename TM>
__global__ __launch_bounds__(kTC) void kFin1(Args<TM> a) {
  const SolveDims D = *a.dims;
  const ApgdState S = a.st[0];
  const long items = (long)D.nc + D.nbil;
  const double* hatV = S.hatSrc >= 0 ? a.ring[S.hatSrc] : a.gHat;
  double acc[kSlots];
  zero(acc);
  GRID_LOOP(q, items, kTC) {
    const bool bil = q >= D.nc;
    const int nr = bil ? 1 : 3;
    double x[3] = {0, 0, 0};
    for (int rr = 0; rr < nr; ++rr) {
      const long o = bil ? BROW(a, q - D.nc) : ROW(a, rr, q);
      const long f = bil ? (q - D.nc) : (D.nbil + 3L * q + rr);
      x[rr] = hatV[o];
      a.gHat[o] = x[rr];
      a.gammaIO[f] = x[rr];
      acc[0] += x[rr] * (double)a.r[o];
    }
    if (!bil) {
      TM F[9];
      ldFrame(a, (int)q, D.geomMode, F);
      storeJ(a, (int)q, F, (TM)x[0], (TM)x[1], (TM)x[2]);
    }
  }
  writePartials<kTC>(acc, a.part1[0]);
}

template <typename TM>
__global__ __launch_bounds__(kTC) void kFin3(Args<TM> a) {
  const SolveDims D = *a.dims;
  const long items = (long)D.nc + D.nbil;
  double acc[kSlots];
  zero(acc);
  GRID_LOOP(q, items, kTC) {
    if (q >= D.nc) {
      acc[1] += a.gHat[BROW(a, q - D.nc)] * (double)bilRow(a, (int)(q - D.nc));
    } else {
      const int c = (int)q;
      TM F[9];
      ldFrame(a, c, D.geomMode, F);
      TM n0, n1, n2;
      contactRows(a, c, F, n0, n1, n2);
      acc[1] += a.gHat[ROW(a, 0, c)] * (double)n0 + a.gHat[ROW(a, 1, c)] * (double)n1 +
                a.gHat[ROW(a, 2, c)] * (double)n2;
    }
  }
  writePartials<kTC>(acc, a.part3[0]);
}

template <typename TM>
__global__ __launch_bounds__(kTD) void kFinDecide(Args<TM> a, int ng) {
  double tot[2 * kSlots];
  reduceAll(a.part1[0], a.part3[0], ng, tot);
  if (threadIdx.x == 0) a.st[0].objective = 0.5 * tot[kSlots + 1] + tot[0];
}

__global__ __launch_bounds__(256) void toFloat(const double* __restrict__ x, float* __restrict__ y, long n) {
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

// ---------------------------------------------------------------------------
// host side

template <typename TM>
struct Plan {
  Args<TM> a;
  int gc, gb;
};

// One chunk of n (even) passes, pass k with parity k & 1, and the decision
// of the last pass.  With `ev` (stream launches only: events recorded inside
// a captured graph cannot be timed), events bracket kernel `prof` of pass k
// (ev[2k], ev[2k+1]); prof 4 times the chunk-end decision (ev[0], ev[1]).
template <typename TM>
void enqueuePasses(const Plan<TM>& P, hipStream_t st, int n, hipEvent_t* ev, int prof) {
  for (int k = 0; k < n; ++k) {
    const int par = k & 1;
    for (int kern = 1; kern <= 3; ++kern) {
      if (ev && prof == kern) DV_HIP(hipEventRecord(ev[2 * k], st));
      if (kern == 1) hipLaunchKernelGGL(kP1<TM>, dim3(P.gc), dim3(kTC), 0, st, P.a, par);
      if (kern == 2) hipLaunchKernelGGL(kBody<TM>, dim3(P.gb), dim3(kTB), 0, st, P.a, 2, par ^ 1);
      if (kern == 3) hipLaunchKernelGGL(kP3<TM>, dim3(P.gc), dim3(kTC), 0, st, P.a, par);
      if (ev && prof == kern) DV_HIP(hipEventRecord(ev[2 * k + 1], st));
    }
  }
  if (ev && prof == 4) DV_HIP(hipEventRecord(ev[0], st));
  hipLaunchKernelGGL(kDecide<TM>, dim3(1), dim3(kTD), 0, st, P.a, (n - 1) & 1, P.gc);
  if (ev && prof == 4) DV_HIP(hipEventRecord(ev[1], st));
}

template <typename TM>
void runSolve(Scene& s, Plan<TM>& P, GraphCache& cache) {
  const hipStream_t st = s.stream;
  // setup (APGD.cu:195-223)
  hipLaunchKernelGGL(kSetup1<TM>, dim3(P.gc), dim3(kTC), 0, st, P.a);
  hipLaunchKernelGGL(kBody<TM>, dim3(P.gb), dim3(kTB), 0, st, P.a, 0, 0);
  hipLaunchKernelGGL(kSetup3<TM>, dim3(P.gc), dim3(kTC), 0, st, P.a);
  hipLaunchKernelGGL(kBody<TM>, dim3(P.gb), dim3(kTB), 0, st, P.a, 1, 0);
  hipLaunchKernelGGL(kSetup5<TM>, dim3(P.gc), dim3(kTC), 0, st, P.a);
  hipLaunchKernelGGL(kSetupDecide<TM>, dim3(1), dim3(kTD), 0, st, P.a, P.gc);
  // iterations: graph chunks until the state reports done
  hipEvent_t* ev = nullptr;
  const int prof = s.profKernel;
  if (prof) {
    while (s.profEv.size() < 2 * (size_t)kPassesPerChunk) {
      hipEvent_t e;
      DV_HIP(hipEventCreate(&e));
      s.profEv.push_back(e);
    }
    ev = s.profEv.data();
  }
  if (!s.noGraph && !cache.matches(&P, sizeof(P))) {
    cache.reset();
    hipGraph_t graph;
    DV_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    enqueuePasses<TM>(P, st, kPassesPerChunk, nullptr, 0);
    DV_HIP(hipStreamEndCapture(st, &graph));
    hipError_t e = hipGraphInstantiate(&cache.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) throw HipError{e, "hipGraphInstantiate(apgd passes)"};
    cache.setKey(&P, sizeof(P));
  }
  const long passCap = (long)std::max(s.maxIt, 1) * 64 + 4096;
  long launched = 0;
  int passesBefore = 0;
  for (;;) {
    // with profiling, the first chunk of the solve is the timed sample
    const bool timed = prof && launched == 0;
    if (s.noGraph || timed) enqueuePasses<TM>(P, st, kPassesPerChunk, timed ? ev : nullptr, prof);
    else DV_HIP(hipGraphLaunch(cache.exec, st));
    launched += kPassesPerChunk;
    DV_HIP(hipMemcpyAsync(s.stateHost, P.a.st, sizeof(ApgdState), hipMemcpyDeviceToHost, st));
    DV_HIP(hipStreamSynchronize(st));
    if (timed) {
      // passes of this chunk that did work are its first (passes - before)
      int ran = std::min(s.stateHost->passes - passesBefore, kPassesPerChunk);
      if (prof == 4) ran = std::min(ran, 1);  // one chunk-end decision per chunk
      for (int k = 0; k < ran; ++k) {
        float ms = 0;
        DV_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
        s.stats.prof_kernel_ms += ms;
        s.stats.prof_kernel_count += 1;
      }
    }
    passesBefore = s.stateHost->passes;
    if (s.stateHost->done) break;
    if (launched > passCap) throw DvError{DV_ERR_DIVERGED, "APGD exceeded its pass budget"};
  }
  // gamma := gammaHat and the objective
  hipLaunchKernelGGL(kFin1<TM>, dim3(P.gc), dim3(kTC), 0, st, P.a);
  hipLaunchKernelGGL(kBody<TM>, dim3(P.gb), dim3(kTB), 0, st, P.a, 3, 0);
  hipLaunchKernelGGL(kFin3<TM>, dim3(P.gc), dim3(kTC), 0, st, P.a);
  hipLaunchKernelGGL(kFinDecide<TM>, dim3(1), dim3(kTD), 0, st, P.a, P.gc);
  DV_HIP(hipMemcpyAsync(s.stateHost, P.a.st, sizeof(ApgdState), hipMemcpyDeviceToHost, st));
  DV_HIP(hipStreamSynchronize(st));
}

}  // namespace

void GraphCache::reset() {
  if (exec) (void)hipGraphExecDestroy(exec);
  exec = nullptr;
  key.clear();
}
bool GraphCache::matches(const void* k, size_t n) const {
  return exec && key.size() == n && std::memcmp(key.data(), k, n) == 0;
}
void GraphCache::setKey(const void* k, size_t n) { key.assign((const char*)k, (const char*)k + n); }

template <typename TM>
void setupPlan(Scene& s, Plan<TM>& P, TM* WM, long RsCap) {
  std::memset(&P, 0, sizeof(P));
  Args<TM>& a = P.a;
  a.dims = s.solveDims;
  a.st = s.state;
  a.cap = s.ncCap;
  a.cA = s.cA.p; a.cB = s.cB.p; a.iStart = s.iStart.p; a.inc = s.inc.p;
  a.bilIncStart = s.bilIncStart.p; a.bilInc = s.bilInc.p; a.bilA = s.bilA.p; a.bilB = s.bilB.p;
  a.heavy = s.heavy.p; a.counters = s.counters.p;
  a.gammaIO = s.gamma.p;
  a.rFlat = s.r.p;
  double* W = s.work.p;
  for (int k = 0; k < 4; ++k) a.ring[k] = W + k * RsCap;
  a.gHat = W + 4 * RsCap;
  for (int k = 0; k < 4; ++k) a.rn[k] = WM + k * RsCap;
  a.r = WM + 4 * RsCap;
  a.delta = WM + 5 * RsCap;
  TM* after = WM + 6 * RsCap;
  // J (per contact) and tmp (per body), 4 components, 32-byte aligned
  uintptr_t p = (reinterpret_cast<uintptr_t>(after) + 31) & ~uintptr_t(31);
  a.J = reinterpret_cast<V4<TM>*>(p);
  a.tmp = a.J + std::max<long>(s.ncCap, 1);
  for (int k = 0; k < 2; ++k) {
    a.part1[k] = s.partials.p + (long)(2 * k) * kSlots * kMaxGridC;
    a.part3[k] = s.partials.p + (long)(2 * k + 1) * kSlots * kMaxGridC;
  }
  const long itemsCap = (long)s.ncCap + s.nbil;
  // one 16-wavefront workgroup per CU (grid-stride beyond): equal work per CU
  P.gc = std::max(1, std::min({blocksFor(itemsCap, kTC), s.numCU, kMaxGridC}));
  P.gb = std::max(1, std::min(blocksFor(s.nb, kTB), 4096));
}

template <typename TM>
size_t workElems(Scene& s, long RsCap) {
  return (size_t)6 * RsCap + 8 + 4 * ((size_t)std::max(s.ncCap, 1) + (size_t)std::max(s.nb, 1)) +
         (sizeof(TM) == 4 ? (size_t)s.nb + 1 : 0) + 16;
}

void solveFused(Scene& s, bool mixed) {
  if (!s.state) {
    DV_HIP(hipMalloc(&s.state, 2 * sizeof(ApgdState)));  // by pass parity
    DV_HIP(hipMalloc(&s.solveDims, sizeof(SolveDims)));
    DV_HIP(hipHostMalloc(&s.stateHost, sizeof(ApgdState), hipHostMallocDefault));
    DV_HIP(hipHostMalloc(&s.dimsHost, sizeof(SolveDims), hipHostMallocDefault));
  }
  s.partials.reserve((size_t)4 * kSlots * kMaxGridC);
  const double cnt = (double)((long)s.nc + s.nbil);
  SolveDims& D = *s.dimsHost;
  D.nc = s.nc; D.nbil = s.nbil; D.nb = s.nb; D.maxIt = s.maxIt;
  D.mu = s.mu; D.tol = s.tol; D.gdiff = 1.0 / (cnt * cnt); D.warm = s.warm ? 1 : 0;
  D.geomMode = s.geomMode;
  DV_HIP(hipMemcpyAsync(s.solveDims, &D, sizeof(SolveDims), hipMemcpyHostToDevice, s.stream));
  // work buffers sized by capacity, so that a captured graph stays valid
  const long RsCap = std::max<long>(3L * s.ncCap + s.nbil, 1);
  s.work.reserve((size_t)5 * RsCap + 16);  // ring[4], gHat: FP64 in both solvers
  if (mixed) {
    s.frame32.reserve((size_t)6 * s.ncCap);  // n, u
    s.work32.reserve(workElems<float>(s, RsCap));
    Plan<float> P;
    setupPlan<float>(s, P, s.work32.p, RsCap);
    float* minv32 = reinterpret_cast<float*>(P.a.tmp + std::max(s.nb, 1));
    if (s.nc)
      hipLaunchKernelGGL(toFloat, dim3(blocksFor(6L * s.ncCap, 256)), dim3(256), 0, s.stream, s.frame.p,
                         s.frame32.p, 6L * s.ncCap);
    hipLaunchKernelGGL(toFloat, dim3(blocksFor(s.nb, 256)), dim3(256), 0, s.stream, s.minv.p, minv32, (long)s.nb);
    P.a.frame = s.frame32.p;
    P.a.minv = minv32;
    runSolve<float>(s, P, s.graphMixed);
  } else {
    s.workM64.reserve(workElems<double>(s, RsCap));
    Plan<double> P;
    setupPlan<double>(s, P, s.workM64.p, RsCap);
    P.a.delta = nullptr;
    P.a.frame = s.frame.p;
    P.a.minv = s.minv.p;
    runSolve<double>(s, P, s.graphFp64);
  }
  const ApgdState& S = *s.stateHost;
  s.stats.iterations = S.it;
  s.stats.backtracks = S.backtracks;
  s.stats.residual = S.best;
  s.stats.objective = S.objective;
  s.stats.passes = S.passes;
  if (s.verbose) printf("  Iterations: %d Residual: %g\n", S.it, S.best);
}

}  // namespace dv
