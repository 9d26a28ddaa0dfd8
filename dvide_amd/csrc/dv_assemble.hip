// Assembly and update kernels of the time step, and the exact-order D / D^T
// products.
//
// D is never materialised as COO: a contact row block is the 3x3 frame
// [n; u; v] (+ for body A, - for body B; System.cu:753-762, 904-909,
// 963-967, 1023-1027) and a bilateral row has +1 / -1 at its DOFs
// (System.cu:622-638).  The products below visit the terms in the order in
// which a row-sorted COO SpMV visits them (D's rows: A's three columns then
// B's; D^T's rows: bilateral rows ascending, then contacts ascending with the
// n, u, v rows in turn), and add them one by one from 0.0 without fused
// multiply-adds, so they round exactly like the oracle's COO SpMV.
#include "dv_internal.h"

namespace dv {
namespace {

constexpr int kT = 256;

// k = M v + h f - h f_elastic  (System.cu:1411-1417, 1502-1514; f_elastic = 0 for bodies)
__global__ __launch_bounds__(kT) void kBuildK(int nb, const double* __restrict__ minv,
                                              const double* __restrict__ v, const double* __restrict__ f,
                                              double* __restrict__ k, double h) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * nb) return;
  double m = minv[i / 3];
  if (m != 0) m = 1.0 / m;
  double kk = m * v[i];
  const double fEl = 0.0;
  k[i] = (h * f[i] + (-h) * fEl) + 1.0 * kk;
}

// y = M^-1 x per DOF (diagonal COO: 0.0 + minv * x)
__global__ __launch_bounds__(kT) void kMinv(int nb, const double* __restrict__ minv,
                                            const double* __restrict__ x, double* __restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * nb) return;
  double s = 0.0;
  s = s + minv[i / 3] * x[i];
  y[i] = s;
}

struct RowArgs {
  int nc, nbil, cap;
  const double* frame;
  const int* cA;
  const int* cB;
  const int* bilA;
  const int* bilB;
};

// y = D x  (+ optional b: y = 1.0*b + D x), one thread per contact / bilateral row
__global__ __launch_bounds__(kT) void kRowsD(RowArgs a, const double* __restrict__ x,
                                             const double* __restrict__ badd, double* __restrict__ y) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.nc + a.nbil) return;
  if (q < a.nbil) {
    double s = 0.0;
    s = s + 1.0 * x[a.bilA[q]];
    if (a.bilB[q] >= 0) s = s + (-1.0) * x[a.bilB[q]];
    y[q] = badd ? 1.0 * badd[q] + s : s;
    return;
  }
  const int c = q - a.nbil;
  const long S = a.cap;
  const int A = a.cA[c], B = a.cB[c];
  const double xa0 = x[3 * A], xa1 = x[3 * A + 1], xa2 = x[3 * A + 2];
  const double xb0 = x[3 * B], xb1 = x[3 * B + 1], xb2 = x[3 * B + 2];
#pragma unroll
  for (int rr = 0; rr < 3; ++rr) {
    const double f0 = a.frame[(3 * rr + 0) * S + c];
    const double f1 = a.frame[(3 * rr + 1) * S + c];
    const double f2 = a.frame[(3 * rr + 2) * S + c];
    double s = 0.0;
    s = s + f0 * xa0;
    s = s + f1 * xa1;
    s = s + f2 * xa2;
    s = s + (-f0) * xb0;
    s = s + (-f1) * xb1;
    s = s + (-f2) * xb2;
    const int row = a.nbil + 3 * c + rr;
    y[row] = badd ? 1.0 * badd[row] + s : s;
  }
}

struct GatherArgs {
  int nb, nbil, cap;
  const double* frame;
  const int* iStart;
  const int* inc;
  const int* bilIncStart;
  const int* bilInc;
  const double* minv;
};

// f = D^T x and tmp = M^-1 f per body, terms in D^T row order.  With
// skipFixed, bodies with M^-1 = 0 get tmp = +0 without the sum (0.0 + 0*f).
__global__ __launch_bounds__(kT) void kGatherDT(GatherArgs a, const double* __restrict__ x,
                                                double* __restrict__ fOut, double* __restrict__ tmpOut,
                                                int skipFixed, int skipHeavy) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.nb) return;
  if (skipHeavy && a.iStart[g + 1] - a.iStart[g] > kHeavyDegree) return;
  const double mi = a.minv[g];
  if (skipFixed && mi == 0) {
    tmpOut[3 * g] = 0.0; tmpOut[3 * g + 1] = 0.0; tmpOut[3 * g + 2] = 0.0;
    return;
  }
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int k = a.bilIncStart[g]; k < a.bilIncStart[g + 1]; ++k) {
    const int e = a.bilInc[k];
    const int row = e >> 3, comp = (e >> 1) & 3;
    const double val = (e & 1) ? -1.0 : 1.0;
    const double t = val * x[row];
    if (comp == 0) s0 = s0 + t;
    else if (comp == 1) s1 = s1 + t;
    else s2 = s2 + t;
  }
  const long S = a.cap;
  for (int k = a.iStart[g]; k < a.iStart[g + 1]; ++k) {
    const int e = a.inc[k];
    const int c = e >> 1;
    const bool sideB = e & 1;
    const long base = a.nbil + 3L * c;
#pragma unroll
    for (int rr = 0; rr < 3; ++rr) {
      const double xr = x[base + rr];
      double v0 = a.frame[(3 * rr + 0) * S + c];
      double v1 = a.frame[(3 * rr + 1) * S + c];
      double v2 = a.frame[(3 * rr + 2) * S + c];
      if (sideB) { v0 = -v0; v1 = -v1; v2 = -v2; }
      s0 = s0 + v0 * xr;
      s1 = s1 + v1 * xr;
      s2 = s2 + v2 * xr;
    }
  }
  if (fOut) { fOut[3 * g] = s0; fOut[3 * g + 1] = s1; fOut[3 * g + 2] = s2; }
  if (tmpOut) {
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    t0 = t0 + mi * s0; t1 = t1 + mi * s1; t2 = t2 + mi * s2;
    tmpOut[3 * g] = t0; tmpOut[3 * g + 1] = t1; tmpOut[3 * g + 2] = t2;
  }
}

// f = D^T x for the bodies of the heavy list (more than kHeavyDegree
// incidences), one workgroup each, FP64 sums in a fixed tree order; used by
// the update of the persistent solvers in place of the sequential sums.
__global__ __launch_bounds__(kT) void kGatherHeavy(GatherArgs a, const int* heavy, const int* counters,
                                                   const double* __restrict__ x, double* __restrict__ fOut) {
  const int nHeavy = counters[2];
  __shared__ double red[kT / 64][3];
  const long S = a.cap;
  for (int hq = blockIdx.x; hq < nHeavy; hq += gridDim.x) {
    const int g = heavy[hq];
    double s0 = 0, s1 = 0, s2 = 0;
    for (int k = a.iStart[g] + threadIdx.x; k < a.iStart[g + 1]; k += kT) {
      const int e = a.inc[k];
      const int c = e >> 1;
      const double sg = (e & 1) ? -1.0 : 1.0;
      const long base = a.nbil + 3L * c;
#pragma unroll
      for (int rr = 0; rr < 3; ++rr) {
        const double xr = sg * x[base + rr];
        s0 += a.frame[(3 * rr + 0) * S + c] * xr;
        s1 += a.frame[(3 * rr + 1) * S + c] * xr;
        s2 += a.frame[(3 * rr + 2) * S + c] * xr;
      }
    }
    if (threadIdx.x == 0) {
      for (int k = a.bilIncStart[g]; k < a.bilIncStart[g + 1]; ++k) {
        const int e = a.bilInc[k];
        const double t = ((e & 1) ? -1.0 : 1.0) * x[e >> 3];
        const int comp = (e >> 1) & 3;
        if (comp == 0) s0 += t; else if (comp == 1) s1 += t; else s2 += t;
      }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      s0 += __shfl_xor(s0, d, 64);
      s1 += __shfl_xor(s1, d, 64);
      s2 += __shfl_xor(s2, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      red[threadIdx.x >> 6][0] = s0; red[threadIdx.x >> 6][1] = s1; red[threadIdx.x >> 6][2] = s2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
      double v = 0;
      for (int w = 0; w < kT / 64; ++w) v += red[w][threadIdx.x];
      fOut[3 * g + threadIdx.x] = v;
    }
    __syncthreads();
  }
}

// b: bilateral rows violation / h (System.cu:1526-1544), contact rows
// (gap/h, 0, 0) (System.cu:1516-1524)
__global__ __launch_bounds__(kT) void kStab(int nc, int nbil, const double* __restrict__ gap,
                                            const int* bilA, const int* bilB, const double* bilVel,
                                            const double* bilT0, const double* bilP0,
                                            const double* __restrict__ p, double h, double time,
                                            double* __restrict__ b) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nc + nbil) return;
  if (q < nbil) {
    double vel = bilVel[q], t0 = bilT0[q];
    if (time < t0) vel = 0;
    double viol = bilB[q] < 0 ? p[bilA[q]] - bilP0[q] - vel * (time - t0)
                              : p[bilA[q]] - p[bilB[q]] - vel * (time - t0);
    b[q] = viol / h;
    return;
  }
  const int c = q - nbil;
  b[nbil + 3 * c] = gap[c] / h;
  b[nbil + 3 * c + 1] = 0;
  b[nbil + 3 * c + 2] = 0;
}

// v = M^-1 (k + f_c), f_c /= h, p += h v  (System.cu:570-581)
__global__ __launch_bounds__(kT) void kUpdate(int nb, const double* __restrict__ minv,
                                              const double* __restrict__ k, double* __restrict__ fc,
                                              double* __restrict__ v, double* __restrict__ p, double h,
                                              int haveRows) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * nb) return;
  const double mi = minv[i / 3];
  double vv;
  if (haveRows) {
    double t = 1.0 * k[i] + 1.0 * fc[i];
    vv = 0.0;
    vv = vv + mi * t;
    fc[i] = (1.0 / h) * fc[i];
  } else {
    vv = 0.0;
    vv = vv + mi * k[i];
    fc[i] = 0.0;
  }
  v[i] = vv;
  p[i] = h * vv + p[i];
}

__global__ __launch_bounds__(kT) void kFillD(double* x, long n, double val) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = val;
}

RowArgs rowArgs(Scene& s) {
  RowArgs a;
  a.nc = s.nc; a.nbil = s.nbil; a.cap = s.ncCap;
  a.frame = s.frame.p; a.cA = s.cA.p; a.cB = s.cB.p;
  a.bilA = s.bilA.p; a.bilB = s.bilB.p;
  return a;
}

GatherArgs gatherArgs(Scene& s) {
  GatherArgs a;
  a.nb = s.nb; a.nbil = s.nbil; a.cap = s.ncCap;
  a.frame = s.frame.p; a.iStart = s.iStart.p; a.inc = s.inc.p;
  a.bilIncStart = s.bilIncStart.p; a.bilInc = s.bilInc.p; a.minv = s.minv.p;
  return a;
}

}  // namespace

void fillDevice(Scene& s, double* x, long n, double v) {
  if (n > 0) hipLaunchKernelGGL(kFillD, dim3(blocksFor(n, kT)), dim3(kT), 0, s.stream, x, n, v);
}

void buildK(Scene& s) {
  hipLaunchKernelGGL(kBuildK, dim3(blocksFor(3L * s.nb, kT)), dim3(kT), 0, s.stream, s.nb, s.minv.p, s.v.p,
                     s.f.p, s.k.p, s.h);
}

void ensureRowBuffers(Scene& s) {
  const int R = s.nbil + 3 * s.nc;
  s.R = R;
  s.r.reserve(std::max(R, 1));
  s.b.reserve(std::max(R, 1));
  // gamma keeps its leading values across steps and zero-fills growth, as
  // thrust::device_vector::resize does (APGD.cu:166)
  s.gamma.reserveKeep(std::max(R, 1), (size_t)s.gammaLen, s.stream);
  if (R > s.gammaLen) fillDevice(s, s.gamma.p + s.gammaLen, R - s.gammaLen, 0.0);
  s.gammaLen = R;
}

// r = D M^-1 k + b  (System.cu:1561-1580)
void buildR(Scene& s) {
  ensureRowBuffers(s);
  const hipStream_t st = s.stream;
  hipLaunchKernelGGL(kMinv, dim3(blocksFor(3L * s.nb, kT)), dim3(kT), 0, st, s.nb, s.minv.p, s.k.p, s.tmp.p);
  const int items = s.nc + s.nbil;
  if (items == 0) return;
  hipLaunchKernelGGL(kStab, dim3(blocksFor(items, kT)), dim3(kT), 0, st, s.nc, s.nbil, s.gap.p, s.bilA.p,
                     s.bilB.p, s.bilVel.p, s.bilT0.p, s.bilP0.p, s.p.p, s.h, s.time, s.b.p);
  hipLaunchKernelGGL(kRowsD, dim3(blocksFor(items, kT)), dim3(kT), 0, st, rowArgs(s), (const double*)s.tmp.p,
                     (const double*)s.b.p, s.r.p);
}

void gatherExact(Scene& s, const double* x, double* fOut, double* tmpOut, bool skipFixed) {
  hipLaunchKernelGGL(kGatherDT, dim3(blocksFor(s.nb, kT)), dim3(kT), 0, s.stream, gatherArgs(s), x, fOut, tmpOut,
                     skipFixed ? 1 : 0, 0);
}

// y = D (M^-1 (D^T x))  (APGD.cu:81-87)
void schurExact(Scene& s, const double* x, double* y) {
  gatherExact(s, x, nullptr, s.tmp.p, true);
  const int items = s.nc + s.nbil;
  if (items)
    hipLaunchKernelGGL(kRowsD, dim3(blocksFor(items, kT)), dim3(kT), 0, s.stream, rowArgs(s), (const double*)s.tmp.p,
                       (const double*)nullptr, y);
}

void minvTimes(Scene& s, const double* x, double* y) {
  if (s.nb) hipLaunchKernelGGL(kMinv, dim3(blocksFor(3L * s.nb, kT)), dim3(kT), 0, s.stream, s.nb, s.minv.p, x, y);
}

void rowsD(Scene& s, const double* x, const double* badd, double* y) {
  const int items = s.nc + s.nbil;
  if (items)
    hipLaunchKernelGGL(kRowsD, dim3(blocksFor(items, kT)), dim3(kT), 0, s.stream, rowArgs(s), x, badd, y);
}

void updateState(Scene& s, bool haveRows, bool exactOrder) {
  if (haveRows) {
    if (exactOrder) {
      gatherExact(s, s.gamma.p, s.fContact.p, nullptr, false);
    } else {
      hipLaunchKernelGGL(kGatherDT, dim3(blocksFor(s.nb, kT)), dim3(kT), 0, s.stream, gatherArgs(s),
                         (const double*)s.gamma.p, s.fContact.p, (double*)nullptr, 0, 1);
      hipLaunchKernelGGL(kGatherHeavy, dim3(64), dim3(kT), 0, s.stream, gatherArgs(s), (const int*)s.heavy.p,
                         (const int*)s.counters.p, (const double*)s.gamma.p, s.fContact.p);
    }
  }
  hipLaunchKernelGGL(kUpdate, dim3(blocksFor(3L * s.nb, kT)), dim3(kT), 0, s.stream, s.nb, s.minv.p, s.k.p,
                     s.fContact.p, s.v.p, s.p.p, s.h, haveRows ? 1 : 0);
}

}  // namespace dv
