// APGD (APGD::solve, APGD.cu:164-361) as a device-resident state machine
// driven by a captured hipGraph: no host round trip per reduction.
//
// One solver pass (one Schur product plus the vector algebra around it) is
// four kernels:
//   P1 (contacts): y and N y from the last two iterates, g = N y + r,
//       gamma_new = Pi(y - t g), the dot products of the objective and
//       restart tests, and J_c = F_c^T gamma_new (the world-frame impulse of
//       contact c on body A);
//   P2 (bodies):   tmp_b = M_b^-1 (sum_A J_c - sum_B J_c + bilateral rows);
//       bodies with many contacts are reduced by a whole workgroup;
//   P3 (contacts): (N gamma_new)_c = F_c (tmp_A - tmp_B), gamma_new . N gamma_new
//       and the residual's infinity norm (getResidual, APGD.cu:90-102);
//   D  (1 block):  fixed-order reduction of the per-workgroup partials, the
//       Lipschitz test (APGD.cu:236-264) and either a backtracking step or
//       the acceptance logic (APGD.cu:266-332) applied to the state.
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
//     are the same vector, so each pass performs one product.
//
// T is the storage/multiply type of D, M^-1 and the iterates: double for the
// FP64 solver, float for the mixed solver.  Dot products accumulate in FP64
// in both.  The mixed solver evaluates the Lipschitz test in the
// cancellation-free form  D'N D >= L |D|^2  (D = gamma_new - y), which equals
// obj1 >= obj2 for g = N y + r and symmetric N, and returns (and measures the
// residual at) the FP64 point Pi(g - gdiff w) of its binary32 iterate g; see
// polish().
#include "dv_geom.h"
#include "dv_internal.h"

#include <algorithm>
#include <cstdio>

namespace dv {

struct ApgdState {
  double L, t, theta, beta, best, objective;
  double sum[kSlots];
  int it, bt, backtracks, restart;
  int cur, prev, nxt, copyPending;
  int hatSrc, done, converged, passes;
};

struct SolveDims {
  int nc, nbil, nb, maxIt;
  double mu, tol, gdiff;
  int warm, pad;
};

namespace {

constexpr int kT = 256;
constexpr int kWaves = kT / 64;
constexpr int kBacktrackCap = 1100;  // L doubles past 1e308 long before this
constexpr int kPassesPerChunk = 16;

template <typename T>
struct Args {
  const SolveDims* dims;
  ApgdState* st;
  const T* frame;
  int cap;
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
  const T* minv;
  const T* r;
  double* gammaIO;      // FP64 flat gamma: previous values in, result out
  T* ring[3];
  T* rn[3];
  T* J;
  T* tmp;
  T* gHat;
  T* gHatN;             // mixed solver: N gammaHat (kept with gammaHat)
  T* delta;
  double* part1;        // kSlots x gmax partials of the contact kernels
  double* part3;
  int gmax;
};

// ---------------------------------------------------------------------------
// block reductions (slot kSlots-1 is a max, the others sums)

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

__device__ void blockReduce(double (&acc)[kSlots], double* out) {
  __shared__ double sh[kWaves][kSlots];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int s = 0; s < kSlots; ++s) {
    double v = (s == kSlots - 1) ? waveMax(acc[s]) : waveSum(acc[s]);
    if (lane == 0) sh[wave][s] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
      double v = sh[0][s];
      for (int w = 1; w < kWaves; ++w) v = (s == kSlots - 1) ? fmax(v, sh[w][s]) : v + sh[w][s];
      out[s] = v;
    }
  }
  __syncthreads();
}

__device__ void writePartials(double (&acc)[kSlots], double* part, int gmax) {
  __shared__ double tot[kSlots];
  blockReduce(acc, tot);
  if (threadIdx.x == 0)
    for (int s = 0; s < kSlots; ++s) part[(long)s * gmax + blockIdx.x] = tot[s];
}

// fixed-order sum of the partials of `ng` workgroups (one block)
__device__ void reducePartials(const double* part, int gmax, int ng, double* out) {
  double a[kSlots];
#pragma unroll
  for (int s = 0; s < kSlots; ++s) {
    double v = 0.0;
    for (int w = threadIdx.x; w < ng; w += kT) {
      const double x = part[(long)s * gmax + w];
      v = (s == kSlots - 1) ? fmax(v, x) : v + x;
    }
    a[s] = v;
  }
  blockReduce(a, out);
}

// ---------------------------------------------------------------------------
// row helpers

template <typename T>
__device__ __forceinline__ void ldFrame(const Args<T>& a, int c, T (&F)[9]) {
  const long S = a.cap;
#pragma unroll
  for (int k = 0; k < 9; ++k) F[k] = a.frame[k * S + c];
}

template <typename T>
__device__ __forceinline__ void storeJ(const Args<T>& a, int c, const T (&F)[9], T x0, T x1, T x2) {
  a.J[3L * c + 0] = F[0] * x0 + F[3] * x1 + F[6] * x2;
  a.J[3L * c + 1] = F[1] * x0 + F[4] * x1 + F[7] * x2;
  a.J[3L * c + 2] = F[2] * x0 + F[5] * x1 + F[8] * x2;
}

template <typename T>
__device__ __forceinline__ T bilRow(const Args<T>& a, int q) {
  T s = a.tmp[a.bilA[q]];
  if (a.bilB[q] >= 0) s -= a.tmp[a.bilB[q]];
  return s;
}

template <typename T>
__device__ __forceinline__ void contactRows(const Args<T>& a, int c, const T (&F)[9], T& n0, T& n1, T& n2) {
  const int A = a.cA[c], B = a.cB[c];
  const T d0 = a.tmp[3 * A] - a.tmp[3 * B];
  const T d1 = a.tmp[3 * A + 1] - a.tmp[3 * B + 1];
  const T d2 = a.tmp[3 * A + 2] - a.tmp[3 * B + 2];
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

// The FP64 point returned for a binary32 iterate g: one projected step of
// length gdiff, g~ = Pi(g - gdiff w).  getResidual divides by
// gdiff = 1/(nc+nbil)^2, so binary32 rounding left in g (a 1e-10 normal
// impulse on a separating contact, say) would dominate its residual; at g~
// those residues are projected away, while g~ - g = O(gdiff |w|) leaves
// w = N g + r unchanged to far below the tolerance.
__device__ __forceinline__ void polish(double& g0, double& g1, double& g2, double w0, double w1, double w2,
                                       double mu, double gdiff) {
  g0 = 1.0 * g0 + (-gdiff) * w0;
  g1 = 1.0 * g1 + (-gdiff) * w1;
  g2 = 1.0 * g2 + (-gdiff) * w2;
  projectCone<double>(g0, g1, g2, mu);
}

__device__ __forceinline__ void zero(double (&acc)[kSlots]) {
#pragma unroll
  for (int s = 0; s < kSlots; ++s) acc[s] = 0.0;
}

// ---------------------------------------------------------------------------
// setup (APGD.cu:195-223)

// gamma_0, gammaHat_0 = 1, d = gamma_0 - gammaHat_0 into ring[1], J = F^T d, |d|^2
template <typename T>
__global__ __launch_bounds__(kT) void kSetup1(Args<T> a) {
  const SolveDims D = *a.dims;
  const long items = (long)D.nbil + D.nc, n3 = 3L * D.nc;
  double acc[kSlots];
  zero(acc);
  T* g0v = a.ring[0];
  T* dv = a.ring[1];
  for (long q = (long)blockIdx.x * kT + threadIdx.x; q < items; q += (long)gridDim.x * kT) {
    const bool bil = q < D.nbil;
    const long f0 = bil ? q : D.nbil + 3L * (q - D.nbil);
    const int nr = bil ? 1 : 3;
    T x[3] = {0, 0, 0};
    for (int rr = 0; rr < nr; ++rr) {
      const long f = f0 + rr;
      // flat entries 0..3nc-1 start as (1,0,0) triples, without the
      // bilateral offset (APGD.cu:116-122); the rest keep their old values
      const double gv = (!D.warm && f < n3) ? ((f % 3 == 0) ? 1.0 : 0.0) : a.gammaIO[f];
      g0v[f] = (T)gv;
      a.gHat[f] = (T)1;
      x[rr] = (T)1 * (T)gv + (T)(-1) * (T)1;
      dv[f] = x[rr];
      acc[0] += (double)x[rr] * (double)x[rr];
    }
    if (!bil) {
      T F[9];
      ldFrame(a, (int)(q - D.nbil), F);
      storeJ(a, (int)(q - D.nbil), F, x[0], x[1], x[2]);
    }
  }
  writePartials(acc, a.part1, a.gmax);
}

// P2: tmp = M^-1 D^T x for non-fixed bodies; X holds the bilateral rows of x.
// which: 0 = ring[1] (d), 1 = ring[0] (gamma_0), 2 = ring[nxt] (gamma_new), 3 = gHat
template <typename T>
__global__ __launch_bounds__(kT) void kBody(Args<T> a, int which) {
  if (which == 2 && a.st->done) return;
  const SolveDims D = *a.dims;
  const T* X = which == 0 ? a.ring[1] : which == 1 ? a.ring[0] : which == 2 ? a.ring[a.st->nxt] : a.gHat;
  for (long g = (long)blockIdx.x * kT + threadIdx.x; g < D.nb; g += (long)gridDim.x * kT) {
    const T mi = a.minv[g];
    const int k0 = a.iStart[g], k1 = a.iStart[g + 1];
    if (mi == (T)0) {
      a.tmp[3 * g] = 0; a.tmp[3 * g + 1] = 0; a.tmp[3 * g + 2] = 0;
      continue;
    }
    if (k1 - k0 > kHeavyDegree) continue;
    T s0 = 0, s1 = 0, s2 = 0;
    for (int k = a.bilIncStart[g]; k < a.bilIncStart[g + 1]; ++k) {
      const int e = a.bilInc[k];
      const T t = (e & 1) ? -X[e >> 3] : X[e >> 3];
      const int comp = (e >> 1) & 3;
      if (comp == 0) s0 += t; else if (comp == 1) s1 += t; else s2 += t;
    }
    for (int k = k0; k < k1; ++k) {
      const int e = a.inc[k];
      const long c = e >> 1;
      const T j0 = a.J[3 * c], j1 = a.J[3 * c + 1], j2 = a.J[3 * c + 2];
      if (e & 1) { s0 -= j0; s1 -= j1; s2 -= j2; }
      else { s0 += j0; s1 += j1; s2 += j2; }
    }
    a.tmp[3 * g] = mi * s0; a.tmp[3 * g + 1] = mi * s1; a.tmp[3 * g + 2] = mi * s2;
  }
  const int nHeavy = a.counters[2];
  for (int hq = blockIdx.x; hq < nHeavy; hq += gridDim.x) {
    const int g = a.heavy[hq];
    const T mi = a.minv[g];
    if (mi == (T)0) continue;
    const int k0 = a.iStart[g], k1 = a.iStart[g + 1];
    double acc[kSlots];
    zero(acc);
    for (int k = k0 + threadIdx.x; k < k1; k += kT) {
      const int e = a.inc[k];
      const long c = e >> 1;
      const double sg = (e & 1) ? -1.0 : 1.0;
      acc[0] += sg * (double)a.J[3 * c];
      acc[1] += sg * (double)a.J[3 * c + 1];
      acc[2] += sg * (double)a.J[3 * c + 2];
    }
    if (threadIdx.x == 0) {
      for (int k = a.bilIncStart[g]; k < a.bilIncStart[g + 1]; ++k) {
        const int e = a.bilInc[k];
        acc[(e >> 1) & 3] += (e & 1) ? -(double)X[e >> 3] : (double)X[e >> 3];
      }
    }
    __shared__ double tot[kSlots];
    blockReduce(acc, tot);
    if (threadIdx.x == 0) {
      a.tmp[3 * g] = mi * (T)tot[0];
      a.tmp[3 * g + 1] = mi * (T)tot[1];
      a.tmp[3 * g + 2] = mi * (T)tot[2];
    }
  }
}

// |N d|^2 into part3[1]; J = F^T gamma_0
template <typename T>
__global__ __launch_bounds__(kT) void kSetup3(Args<T> a) {
  const SolveDims D = *a.dims;
  const long items = (long)D.nbil + D.nc;
  double acc[kSlots];
  zero(acc);
  for (long q = (long)blockIdx.x * kT + threadIdx.x; q < items; q += (long)gridDim.x * kT) {
    if (q < D.nbil) {
      const double v = (double)bilRow(a, (int)q);
      acc[1] += v * v;
    } else {
      const int c = (int)(q - D.nbil);
      T F[9];
      ldFrame(a, c, F);
      T n0, n1, n2;
      contactRows(a, c, F, n0, n1, n2);
      acc[1] += (double)n0 * (double)n0 + (double)n1 * (double)n1 + (double)n2 * (double)n2;
      const long f = D.nbil + 3L * c;
      storeJ(a, c, F, a.ring[0][f], a.ring[0][f + 1], a.ring[0][f + 2]);
    }
  }
  writePartials(acc, a.part3, a.gmax);
}

// N gamma_0 into rn[0]
template <typename T>
__global__ __launch_bounds__(kT) void kSetup5(Args<T> a) {
  const SolveDims D = *a.dims;
  const long items = (long)D.nbil + D.nc;
  for (long q = (long)blockIdx.x * kT + threadIdx.x; q < items; q += (long)gridDim.x * kT) {
    if (q < D.nbil) {
      a.rn[0][q] = bilRow(a, (int)q);
    } else {
      const int c = (int)(q - D.nbil);
      T F[9];
      ldFrame(a, c, F);
      T n0, n1, n2;
      contactRows(a, c, F, n0, n1, n2);
      const long f = D.nbil + 3L * c;
      a.rn[0][f] = n0; a.rn[0][f + 1] = n1; a.rn[0][f + 2] = n2;
    }
  }
}

// L_0 = |N d| / |d|, t_0 = 1/L_0 and the initial state (APGD.cu:208-223)
template <typename T>
__global__ __launch_bounds__(kT) void kSetupDecide(Args<T> a, int ng) {
  __shared__ double s1[kSlots], s3[kSlots];
  reducePartials(a.part1, a.gmax, ng, s1);
  reducePartials(a.part3, a.gmax, ng, s3);
  if (threadIdx.x == 0) {
    ApgdState S;
    double L = sqrt(s1[0]);
    L = sqrt(s3[1]) / L;
    S.L = L;
    S.t = 1.0 / L;
    S.theta = 1.0;
    S.beta = 0.0;
    S.best = 10e30;
    S.objective = 0;
    for (int k = 0; k < kSlots; ++k) S.sum[k] = 0;
    S.it = 0; S.bt = 0; S.backtracks = 0; S.restart = 1;
    S.cur = 0; S.prev = 0; S.nxt = 1; S.copyPending = 0; S.hatSrc = -1;
    S.done = a.dims->maxIt <= 0 ? 1 : 0;
    S.converged = 0;
    S.passes = 0;
    *a.st = S;
  }
}

// ---------------------------------------------------------------------------
// one pass

template <typename T>
__global__ __launch_bounds__(kT) void kP1(Args<T> a) {
  constexpr bool kMixed = sizeof(T) < sizeof(double);
  const ApgdState S = *a.st;
  if (S.done) return;
  const SolveDims D = *a.dims;
  const long items = (long)D.nbil + D.nc;
  const T* Gk = a.ring[S.cur];
  const T* Gkm1 = a.ring[S.prev];
  const T* Nk = a.rn[S.cur];
  const T* Nkm1 = a.rn[S.prev];
  T* Gn = a.ring[S.nxt];
  const T bp = (T)(1.0 + S.beta), bm = (T)(-S.beta);
  const T tT = (T)S.t;
  const T mu = (T)D.mu;
  const bool doCopy = S.copyPending && S.bt == 0;
  const T* hatSrc = doCopy ? a.ring[S.hatSrc] : nullptr;
  const T* hatSrcN = doCopy ? a.rn[S.hatSrc] : nullptr;
  double acc[kSlots];
  zero(acc);
  for (long q = (long)blockIdx.x * kT + threadIdx.x; q < items; q += (long)gridDim.x * kT) {
    const bool bil = q < D.nbil;
    const int c = bil ? -1 : (int)(q - D.nbil);
    const long f0 = bil ? q : D.nbil + 3L * c;
    const int nr = bil ? 1 : 3;
    T y[3] = {0, 0, 0}, ny[3] = {0, 0, 0}, gk[3] = {0, 0, 0}, gg[3] = {0, 0, 0}, gn[3] = {0, 0, 0};
#pragma unroll
    for (int rr = 0; rr < 3; ++rr) {
      if (rr < nr) {
        const long f = f0 + rr;
        gk[rr] = Gk[f];
        if (S.restart) {
          y[rr] = gk[rr];
          ny[rr] = Nk[f];
        } else {
          y[rr] = bp * gk[rr] + bm * Gkm1[f];
          ny[rr] = bp * Nk[f] + bm * Nkm1[f];
        }
        gg[rr] = (T)1 * ny[rr] + (T)1 * a.r[f];
        gn[rr] = (T)1 * y[rr] + (-tT) * gg[rr];
        if (doCopy) {
          a.gHat[f] = hatSrc[f];
          if (kMixed) a.gHatN[f] = hatSrcN[f];
        }
      }
    }
    if (!bil) projectCone<T>(gn[0], gn[1], gn[2], mu);
#pragma unroll
    for (int rr = 0; rr < 3; ++rr) {
      if (rr < nr) {
        const long f = f0 + rr;
        Gn[f] = gn[rr];
        const double dy = (double)gn[rr] - (double)y[rr];
        if (kMixed) {
          a.delta[f] = gn[rr] - y[rr];
          acc[1] += dy * (double)ny[rr];
        } else {
          acc[0] += (double)gn[rr] * (double)a.r[f];
          acc[1] += (double)y[rr] * (double)ny[rr];
          acc[2] += (double)y[rr] * (double)a.r[f];
          acc[3] += (double)gg[rr] * dy;
        }
        acc[4] += dy * dy;
        acc[5] += (double)gg[rr] * ((double)gn[rr] - (double)gk[rr]);
      }
    }
    if (!bil) {
      T F[9];
      ldFrame(a, c, F);
      storeJ(a, c, F, gn[0], gn[1], gn[2]);
    }
  }
  writePartials(acc, a.part1, a.gmax);
}

template <typename T>
__global__ __launch_bounds__(kT) void kP3(Args<T> a) {
  constexpr bool kMixed = sizeof(T) < sizeof(double);
  const ApgdState S = *a.st;
  if (S.done) return;
  const SolveDims D = *a.dims;
  const long items = (long)D.nbil + D.nc;
  const T* Gn = a.ring[S.nxt];
  T* Nn = a.rn[S.nxt];
  double acc[kSlots];
  zero(acc);
  for (long q = (long)blockIdx.x * kT + threadIdx.x; q < items; q += (long)gridDim.x * kT) {
    if (q < D.nbil) {
      const T nv = bilRow(a, (int)q);
      Nn[q] = nv;
      double gv = (double)Gn[q];
      acc[6] += (kMixed ? (double)a.delta[q] : gv) * (double)nv;
      const double w = (double)nv + (double)a.r[q];
      if (kMixed && D.nc) gv = 1.0 * gv + (-D.gdiff) * w;
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
      const int c = (int)(q - D.nbil);
      T F[9];
      ldFrame(a, c, F);
      T n0, n1, n2;
      contactRows(a, c, F, n0, n1, n2);
      const long f = D.nbil + 3L * c;
      Nn[f] = n0; Nn[f + 1] = n1; Nn[f + 2] = n2;
      double g0 = (double)Gn[f], g1 = (double)Gn[f + 1], g2 = (double)Gn[f + 2];
      const double w0 = (double)n0 + (double)a.r[f];
      const double w1 = (double)n1 + (double)a.r[f + 1];
      const double w2 = (double)n2 + (double)a.r[f + 2];
      if (kMixed) {
        acc[6] += (double)a.delta[f] * (double)n0 + (double)a.delta[f + 1] * (double)n1 +
                  (double)a.delta[f + 2] * (double)n2;
        polish(g0, g1, g2, w0, w1, w2, D.mu, D.gdiff);
      } else {
        acc[6] += g0 * (double)n0 + g1 * (double)n1 + g2 * (double)n2;
      }
      acc[7] = fmax(acc[7], residualContact(g0, g1, g2, w0, w1, w2, D.mu, D.gdiff));
    }
  }
  writePartials(acc, a.part3, a.gmax);
}

// Lipschitz test, backtracking or acceptance (APGD.cu:236-332)
template <typename T>
__global__ __launch_bounds__(kT) void kDecide(Args<T> a, int ng) {
  constexpr bool kMixed = sizeof(T) < sizeof(double);
  if (a.st->done) return;
  __shared__ double s1[kSlots], s3[kSlots];
  reducePartials(a.part1, a.gmax, ng, s1);
  reducePartials(a.part3, a.gmax, ng, s3);
  if (threadIdx.x != 0) return;
  ApgdState S = *a.st;
  const SolveDims D = *a.dims;
  double sum[kSlots];
  for (int k = 0; k < 6; ++k) sum[k] = s1[k];
  sum[6] = s3[6];
  sum[7] = s3[7];
  for (int k = 0; k < kSlots; ++k) S.sum[k] = sum[k];
  S.passes++;
  if (S.bt == 0) S.copyPending = 0;  // this pass's P1 performed the pending copy
  const double L = S.L;
  bool backtrack;
  if (kMixed) {
    backtrack = (sum[6] - sum[1]) >= L * sum[4];
  } else {
    const double obj1 = 0.5 * sum[6] + sum[0];
    const double nrm = sqrt(sum[4]);
    double obj2 = 0.5 * sum[1] + sum[2];
    obj2 += sum[3] + 0.5 * L * (nrm * nrm);
    backtrack = obj1 >= obj2;
  }
  if (backtrack && S.bt < kBacktrackCap) {
    S.bt++;
    S.backtracks++;
    S.L = 2.0 * L;
    S.t = 1.0 / S.L;
    *a.st = S;
    return;
  }
  const double theta = S.theta;
  double thetaNew = (-(theta * theta) + theta * sqrt(theta * theta + 4.0)) / 2.0;
  S.beta = theta * (1.0 - theta) / (theta * theta + thetaNew);
  const double res = sum[7];
  if (res < S.best) {
    S.best = res;
    S.copyPending = 1;
    S.hatSrc = S.nxt;
  }
  if (S.best < D.tol) {
    S.done = 1;
    S.converged = 1;
    *a.st = S;
    return;
  }
  if (sum[5] > 0) {
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
  S.nxt = 3 - S.prev - S.cur;
  S.it++;
  S.bt = 0;
  if (S.it >= D.maxIt) S.done = 1;
  *a.st = S;
}

// ---------------------------------------------------------------------------
// gamma := gammaHat (APGD.cu:358) and objectiveCCP (APGD.cu:353-354)

template <typename T>
__global__ __launch_bounds__(kT) void kFin1(Args<T> a) {
  constexpr bool kMixed = sizeof(T) < sizeof(double);
  const SolveDims D = *a.dims;
  const ApgdState S = *a.st;
  const long items = (long)D.nbil + D.nc;
  const T* hatV = S.copyPending ? a.ring[S.hatSrc] : a.gHat;
  const T* hatN = kMixed ? (S.copyPending ? a.rn[S.hatSrc] : a.gHatN) : nullptr;
  double acc[kSlots];
  zero(acc);
  for (long q = (long)blockIdx.x * kT + threadIdx.x; q < items; q += (long)gridDim.x * kT) {
    const bool bil = q < D.nbil;
    const long f0 = bil ? q : D.nbil + 3L * (q - D.nbil);
    const int nr = bil ? 1 : 3;
    T x[3] = {0, 0, 0};
    double xd[3] = {0, 0, 0}, w[3] = {0, 0, 0};
    for (int rr = 0; rr < nr; ++rr) {
      x[rr] = hatV[f0 + rr];
      a.gHat[f0 + rr] = x[rr];
      xd[rr] = (double)x[rr];
      acc[0] += (double)x[rr] * (double)a.r[f0 + rr];
      if (kMixed) w[rr] = (double)hatN[f0 + rr] + (double)a.r[f0 + rr];
    }
    // the mixed solver returns the FP64 point its residual was measured at
    if (kMixed && D.nc && S.best < 1e30) {  // (no accepted iterate: gammaHat = 1)
      if (bil) xd[0] = 1.0 * xd[0] + (-D.gdiff) * w[0];
      else polish(xd[0], xd[1], xd[2], w[0], w[1], w[2], D.mu, D.gdiff);
    }
    for (int rr = 0; rr < nr; ++rr) a.gammaIO[f0 + rr] = xd[rr];
    if (!bil) {
      T F[9];
      ldFrame(a, (int)(q - D.nbil), F);
      storeJ(a, (int)(q - D.nbil), F, x[0], x[1], x[2]);
    }
  }
  writePartials(acc, a.part1, a.gmax);
}

template <typename T>
__global__ __launch_bounds__(kT) void kFin3(Args<T> a) {
  const SolveDims D = *a.dims;
  const long items = (long)D.nbil + D.nc;
  double acc[kSlots];
  zero(acc);
  for (long q = (long)blockIdx.x * kT + threadIdx.x; q < items; q += (long)gridDim.x * kT) {
    if (q < D.nbil) {
      acc[1] += (double)a.gHat[q] * (double)bilRow(a, (int)q);
    } else {
      const int c = (int)(q - D.nbil);
      T F[9];
      ldFrame(a, c, F);
      T n0, n1, n2;
      contactRows(a, c, F, n0, n1, n2);
      const long f = D.nbil + 3L * c;
      acc[1] += (double)a.gHat[f] * (double)n0 + (double)a.gHat[f + 1] * (double)n1 +
                (double)a.gHat[f + 2] * (double)n2;
    }
  }
  writePartials(acc, a.part3, a.gmax);
}

template <typename T>
__global__ __launch_bounds__(kT) void kFinDecide(Args<T> a, int ng) {
  __shared__ double s1[kSlots], s3[kSlots];
  reducePartials(a.part1, a.gmax, ng, s1);
  reducePartials(a.part3, a.gmax, ng, s3);
  if (threadIdx.x == 0) a.st->objective = 0.5 * s3[1] + s1[0];
}

__global__ __launch_bounds__(kT) void toFloat(const double* __restrict__ x, float* __restrict__ y, long n) {
  long i = (long)blockIdx.x * kT + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

// ---------------------------------------------------------------------------
// host side

template <typename T>
struct Plan {
  Args<T> a;
  int gc, gb;
};

template <typename T>
void capturePasses(Scene& s, const Plan<T>& P, hipGraphExec_t* exec) {
  hipGraph_t graph;
  DV_HIP(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < kPassesPerChunk; ++k) {
    hipLaunchKernelGGL(kP1<T>, dim3(P.gc), dim3(kT), 0, s.stream, P.a);
    hipLaunchKernelGGL(kBody<T>, dim3(P.gb), dim3(kT), 0, s.stream, P.a, 2);
    hipLaunchKernelGGL(kP3<T>, dim3(P.gc), dim3(kT), 0, s.stream, P.a);
    hipLaunchKernelGGL(kDecide<T>, dim3(1), dim3(kT), 0, s.stream, P.a, P.gc);
  }
  DV_HIP(hipStreamEndCapture(s.stream, &graph));
  hipError_t e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) throw HipError{e, "hipGraphInstantiate(apgd passes)"};
}

template <typename T>
void runSolve(Scene& s, Plan<T>& P, GraphCache& cache) {
  const hipStream_t st = s.stream;
  // setup (APGD.cu:195-223)
  hipLaunchKernelGGL(kSetup1<T>, dim3(P.gc), dim3(kT), 0, st, P.a);
  hipLaunchKernelGGL(kBody<T>, dim3(P.gb), dim3(kT), 0, st, P.a, 0);
  hipLaunchKernelGGL(kSetup3<T>, dim3(P.gc), dim3(kT), 0, st, P.a);
  hipLaunchKernelGGL(kBody<T>, dim3(P.gb), dim3(kT), 0, st, P.a, 1);
  hipLaunchKernelGGL(kSetup5<T>, dim3(P.gc), dim3(kT), 0, st, P.a);
  hipLaunchKernelGGL(kSetupDecide<T>, dim3(1), dim3(kT), 0, st, P.a, P.gc);
  // iterations: graph chunks until the state reports done
  if (!s.noGraph && !cache.matches(&P, sizeof(P))) {
    cache.reset();
    capturePasses<T>(s, P, &cache.exec);
    cache.setKey(&P, sizeof(P));
  }
  const long passCap = (long)std::max(s.maxIt, 1) * 64 + 4096;
  long launched = 0;
  for (;;) {
    if (s.noGraph) {
      for (int k = 0; k < kPassesPerChunk; ++k) {
        hipLaunchKernelGGL(kP1<T>, dim3(P.gc), dim3(kT), 0, st, P.a);
        hipLaunchKernelGGL(kBody<T>, dim3(P.gb), dim3(kT), 0, st, P.a, 2);
        hipLaunchKernelGGL(kP3<T>, dim3(P.gc), dim3(kT), 0, st, P.a);
        hipLaunchKernelGGL(kDecide<T>, dim3(1), dim3(kT), 0, st, P.a, P.gc);
      }
    } else {
      DV_HIP(hipGraphLaunch(cache.exec, st));
    }
    launched += kPassesPerChunk;
    DV_HIP(hipMemcpyAsync(&s.stateHost->done, &P.a.st->done, sizeof(int), hipMemcpyDeviceToHost, st));
    DV_HIP(hipStreamSynchronize(st));
    if (s.stateHost->done) break;
    if (launched > passCap) throw DvError{DV_ERR_DIVERGED, "APGD exceeded its pass budget"};
  }
  // gamma := gammaHat and the objective
  hipLaunchKernelGGL(kFin1<T>, dim3(P.gc), dim3(kT), 0, st, P.a);
  hipLaunchKernelGGL(kBody<T>, dim3(P.gb), dim3(kT), 0, st, P.a, 3);
  hipLaunchKernelGGL(kFin3<T>, dim3(P.gc), dim3(kT), 0, st, P.a);
  hipLaunchKernelGGL(kFinDecide<T>, dim3(1), dim3(kT), 0, st, P.a, P.gc);
  DV_HIP(hipMemcpyAsync(s.stateHost, P.a.st, sizeof(ApgdState), hipMemcpyDeviceToHost, st));
  DV_HIP(hipStreamSynchronize(st));
}

template <typename T>
void fillCommon(Scene& s, Args<T>& a) {
  a.dims = s.solveDims;
  a.st = s.state;
  a.cap = s.ncCap;
  a.cA = s.cA.p; a.cB = s.cB.p; a.iStart = s.iStart.p; a.inc = s.inc.p;
  a.bilIncStart = s.bilIncStart.p; a.bilInc = s.bilInc.p; a.bilA = s.bilA.p; a.bilB = s.bilB.p;
  a.heavy = s.heavy.p; a.counters = s.counters.p;
  a.gammaIO = s.gamma.p;
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

void solveFused(Scene& s, bool mixed) {
  const long R = s.R;
  const long Js = std::max<long>(3L * s.ncCap, 3);
  const long Bs = std::max<long>(3L * s.nb, 3);
  if (!s.state) {
    DV_HIP(hipMalloc(&s.state, sizeof(ApgdState)));
    DV_HIP(hipMalloc(&s.solveDims, sizeof(SolveDims)));
    DV_HIP(hipHostMalloc(&s.stateHost, sizeof(ApgdState), hipHostMallocDefault));
    DV_HIP(hipHostMalloc(&s.dimsHost, sizeof(SolveDims), hipHostMallocDefault));
  }
  const long itemsCap = (long)s.ncCap + s.nbil;
  const int gc = std::max(1, std::min(blocksFor(itemsCap, kT), 2048));
  const int gb = std::max(1, std::min(blocksFor(s.nb, kT), 2048));
  s.partials.reserve((size_t)2 * kSlots * 2048);
  const double cnt = (double)((long)s.nc + s.nbil);
  SolveDims& D = *s.dimsHost;
  D.nc = s.nc; D.nbil = s.nbil; D.nb = s.nb; D.maxIt = s.maxIt;
  D.mu = s.mu; D.tol = s.tol; D.gdiff = 1.0 / (cnt * cnt); D.warm = s.warm ? 1 : 0; D.pad = 0;
  DV_HIP(hipMemcpyAsync(s.solveDims, &D, sizeof(SolveDims), hipMemcpyHostToDevice, s.stream));
  // work buffers sized by capacity so that a captured graph stays valid
  const long RsCap = std::max<long>(3L * s.ncCap + s.nbil, 1);
  if (mixed) {
    s.frame32.reserve((size_t)9 * s.ncCap);
    s.work32.reserve((size_t)10 * RsCap + Js + Bs + Bs / 3 + 16);
    float* W = s.work32.p;
    Plan<float> P;
    std::memset(&P, 0, sizeof(P));
    for (int k = 0; k < 3; ++k) P.a.ring[k] = W + k * RsCap;
    for (int k = 0; k < 3; ++k) P.a.rn[k] = W + (3 + k) * RsCap;
    P.a.gHat = W + 6 * RsCap;
    P.a.delta = W + 7 * RsCap;
    float* r32 = W + 8 * RsCap;
    P.a.gHatN = W + 9 * RsCap;
    P.a.J = W + 10 * RsCap;
    P.a.tmp = P.a.J + Js;
    float* minv32 = P.a.tmp + Bs;
    if (s.nc)
      hipLaunchKernelGGL(toFloat, dim3(blocksFor(9L * s.ncCap, kT)), dim3(kT), 0, s.stream, s.frame.p, s.frame32.p,
                         9L * s.ncCap);
    if (R) hipLaunchKernelGGL(toFloat, dim3(blocksFor(R, kT)), dim3(kT), 0, s.stream, s.r.p, r32, R);
    hipLaunchKernelGGL(toFloat, dim3(blocksFor(s.nb, kT)), dim3(kT), 0, s.stream, s.minv.p, minv32, (long)s.nb);
    fillCommon(s, P.a);
    P.a.frame = s.frame32.p;
    P.a.minv = minv32;
    P.a.r = r32;
    P.a.part1 = s.partials.p;
    P.a.part3 = s.partials.p + (long)kSlots * 2048;
    P.a.gmax = 2048;
    P.gc = gc;
    P.gb = gb;
    runSolve<float>(s, P, s.graphMixed);
  } else {
    s.work.reserve((size_t)7 * RsCap + Js + Bs + 16);
    double* W = s.work.p;
    Plan<double> P;
    std::memset(&P, 0, sizeof(P));
    for (int k = 0; k < 3; ++k) P.a.ring[k] = W + k * RsCap;
    for (int k = 0; k < 3; ++k) P.a.rn[k] = W + (3 + k) * RsCap;
    P.a.gHat = W + 6 * RsCap;
    P.a.gHatN = nullptr;
    P.a.delta = nullptr;
    P.a.J = W + 7 * RsCap;
    P.a.tmp = P.a.J + Js;
    fillCommon(s, P.a);
    P.a.frame = s.frame.p;
    P.a.minv = s.minv.p;
    P.a.r = s.r.p;
    P.a.part1 = s.partials.p;
    P.a.part3 = s.partials.p + (long)kSlots * 2048;
    P.a.gmax = 2048;
    P.gc = gc;
    P.gb = gb;
    runSolve<double>(s, P, s.graphFp64);
  }
  const ApgdState& S = *s.stateHost;
  s.stats.iterations = S.it;
  s.stats.backtracks = S.backtracks;
  s.stats.residual = S.best;
  s.stats.objective = S.objective;
  if (s.verbose) printf("  Iterations: %d Residual: %g\n", S.it, S.best);
}

}  // namespace dv
