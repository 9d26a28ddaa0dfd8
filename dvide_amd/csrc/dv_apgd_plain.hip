// The plain FP64 APGD path: APGD::solve (APGD.cu:164-361) driven from the
// host exactly as the reference drives it -- one kernel per CUSP operation,
// four Schur products per iteration (N y twice, N gamma_new, and the one
// inside getResidual), and every dot product / norm read back to the host.
//
// It is the FP64 baseline of the GPU path.  Its products use the exact-order
// D / D^T kernels and its reductions the fixed CR256 tree (blocks of 256
// consecutive terms reduced with strides 128..1, block results summed left to
// right on the host), so every scalar it computes rounds like the oracle's.
#include "dv_geom.h"
#include "dv_internal.h"

#include <cmath>
#include <cstdio>

namespace dv {
namespace {

constexpr int kT = 256;

__global__ __launch_bounds__(kT) void kAxpby(const double* __restrict__ x, const double* __restrict__ y,
                                             double* __restrict__ z, double a, double b, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = a * x[i] + b * y[i];
}

// updateAntiRelaxationVector (APGD.cu:104-114): ar[3c] = mu |(s_u, s_v)| of
// s = D v, ar[3c+1] = ar[3c+2] = 0, indexed from 0 without the bilateral
// offset as the reference indexes it
__global__ __launch_bounds__(kT) void kAntiRelax(const double* __restrict__ sv, double* __restrict__ ar, int nc,
                                                 double mu) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const double a = sv[3 * c + 1], b = sv[3 * c + 2];
  ar[3 * c] = sqrt(a * a + b * b) * mu;
  ar[3 * c + 1] = 0.0;
  ar[3 * c + 2] = 0.0;
}

__global__ __launch_bounds__(kT) void kInitGamma(double* g, int nc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  g[3 * i] = 1.0;
  g[3 * i + 1] = 0.0;
  g[3 * i + 2] = 0.0;
}

__global__ __launch_bounds__(kT) void kProject(double* g, int off, int nc, double mu) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double* q = g + off + 3L * c;
  double a = q[0], b = q[1], d = q[2];
  projectCone<double>(a, b, d, mu);
  q[0] = a; q[1] = b; q[2] = d;
}

// CR256 partial sums of x.y (y == nullptr: x.x)
__global__ __launch_bounds__(kT) void kDot256(const double* __restrict__ x, const double* __restrict__ y,
                                              long n, double* __restrict__ part) {
  __shared__ double t[kT];
  long i = (long)blockIdx.x * kT + threadIdx.x;
  double v = 0.0;
  if (i < n) v = y ? x[i] * y[i] : x[i] * x[i];
  t[threadIdx.x] = v;
  __syncthreads();
  for (int st = kT / 2; st >= 1; st >>= 1) {
    if ((int)threadIdx.x < st) t[threadIdx.x] = t[threadIdx.x] + t[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = t[0];
}

__global__ __launch_bounds__(kT) void kMaxAbs(const double* __restrict__ x, long n, double* __restrict__ part) {
  __shared__ double t[kT];
  long i = (long)blockIdx.x * kT + threadIdx.x;
  const double a = i < n ? fabs(x[i]) : 0.0;
  t[threadIdx.x] = a == a ? a : INFINITY;  // fmax drops NaN; +inf is kept
  __syncthreads();
  for (int st = kT / 2; st >= 1; st >>= 1) {
    if ((int)threadIdx.x < st) t[threadIdx.x] = fmax(t[threadIdx.x], t[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = t[0];
}

struct Plain {
  Scene& s;
  long R;
  int nblk;
  double* part;     // device partials
  double* hpart;    // pinned host partials
  Plain(Scene& sc, long r, double* dpart, double* hp) : s(sc), R(r), nblk(blocksFor(r, kT)), part(dpart), hpart(hp) {}

  void axpby(const double* x, const double* y, double* z, double a, double b) {
    if (R) hipLaunchKernelGGL(kAxpby, dim3(nblk), dim3(kT), 0, s.stream, x, y, z, a, b, R);
  }
  void copy(const double* x, double* y) {
    if (R) DV_HIP(hipMemcpyAsync(y, x, sizeof(double) * R, hipMemcpyDeviceToDevice, s.stream));
  }
  double reduceSum() {
    DV_HIP(hipMemcpyAsync(hpart, part, sizeof(double) * nblk, hipMemcpyDeviceToHost, s.stream));
    DV_HIP(hipStreamSynchronize(s.stream));
    double acc = 0;
    for (int i = 0; i < nblk; ++i) acc += hpart[i];
    return acc;
  }
  double dot(const double* x, const double* y) {
    if (!R) return 0.0;
    hipLaunchKernelGGL(kDot256, dim3(nblk), dim3(kT), 0, s.stream, x, y, R, part);
    return reduceSum();
  }
  double nrm2(const double* x) {
    if (!R) return 0.0;
    hipLaunchKernelGGL(kDot256, dim3(nblk), dim3(kT), 0, s.stream, x, (const double*)nullptr, R, part);
    return std::sqrt(reduceSum());
  }
  double nrmmax(const double* x) {
    if (!R) return 0.0;
    hipLaunchKernelGGL(kMaxAbs, dim3(nblk), dim3(kT), 0, s.stream, x, R, part);
    DV_HIP(hipMemcpyAsync(hpart, part, sizeof(double) * nblk, hipMemcpyDeviceToHost, s.stream));
    DV_HIP(hipStreamSynchronize(s.stream));
    double m = 0;
    for (int i = 0; i < nblk; ++i) m = std::max(m, hpart[i]);
    return m;
  }
  void project(double* x) {
    if (s.nc)
      hipLaunchKernelGGL(kProject, dim3(blocksFor(s.nc, kT)), dim3(kT), 0, s.stream, x, s.nbil, s.nc, s.mu);
  }
  void schur(const double* x, double* y) { schurExact(s, x, y); }
};

}  // namespace

void solvePlain(Scene& s) {
  const long R = s.R;
  // work vectors: gammaHat, gammaNew, g, y, yNew, t, Ny, Ng, the
  // anti-relaxation vector
  s.work.reserve((size_t)9 * std::max<long>(R, 1));
  double* W = s.work.p;
  const long S = std::max<long>(R, 1);
  double *gHat = W, *gNew = W + S, *g = W + 2 * S, *y = W + 3 * S, *yNew = W + 4 * S, *t = W + 5 * S,
         *Ny = W + 6 * S, *Ng = W + 7 * S, *ar = W + 8 * S;
  const int nblk = blocksFor(std::max<long>(R, 1), kT);
  s.partials.reserve(nblk);
  double* hp = nullptr;
  DV_HIP(hipHostMalloc(&hp, sizeof(double) * nblk, hipHostMallocDefault));
  struct HostFree {
    double* p;
    ~HostFree() { (void)hipHostFree(p); }
  } guard{hp};
  Plain P(s, R, s.partials.p, hp);
  double* gamma = s.gamma.p;

  if (!s.warm && s.nc)
    hipLaunchKernelGGL(kInitGamma, dim3(blocksFor(s.nc, kT)), dim3(kT), 0, s.stream, gamma, s.nc);
  fillDevice(s, gHat, R, 1.0);
  if (s.antiRelax) fillDevice(s, ar, R, 0.0);  // APGD.cu:189
  P.copy(gamma, y);
  double theta = 1.0, thetaNew = 1.0, beta = 0, obj1 = 0, obj2 = 0, best = 10e30;
  P.axpby(gamma, gHat, t, 1.0, -1.0);
  double L = P.nrm2(t);
  P.schur(t, Ng);
  L = P.nrm2(Ng) / L;
  double tt = 1.0 / L;
  const double cnt = (double)((long)s.nc + s.nbil);
  const double gdiff = 1.0 / (cnt * cnt);
  int backtracks = 0;
  int it;
  for (it = 0; it < s.maxIt; ++it) {
    P.schur(y, Ny);
    P.axpby(Ny, s.r.p, g, 1.0, 1.0);
    P.axpby(y, g, gNew, 1.0, -tt);
    P.project(gNew);
    P.schur(gNew, Ng);
    obj1 = 0.5 * P.dot(gNew, Ng) + P.dot(gNew, s.r.p);
    P.schur(y, Ny);
    obj2 = 0.5 * P.dot(y, Ny) + P.dot(y, s.r.p);
    P.axpby(gNew, y, t, 1.0, -1.0);
    double nr = P.nrm2(t);
    obj2 += P.dot(g, t) + 0.5 * L * (nr * nr);
    while (obj1 >= obj2) {
      ++backtracks;
      L = 2.0 * L;
      tt = 1.0 / L;
      P.axpby(y, g, gNew, 1.0, -tt);
      P.project(gNew);
      P.schur(gNew, Ng);
      obj1 = 0.5 * P.dot(gNew, Ng) + P.dot(gNew, s.r.p);
      P.schur(y, Ny);
      obj2 = 0.5 * P.dot(y, Ny) + P.dot(y, s.r.p);
      P.axpby(gNew, y, t, 1.0, -1.0);
      nr = P.nrm2(t);
      obj2 += P.dot(g, t) + 0.5 * L * (nr * nr);
    }
    thetaNew = (-(theta * theta) + theta * std::sqrt(theta * theta + 4.0)) / 2.0;
    beta = theta * (1.0 - theta) / (theta * theta + thetaNew);
    P.axpby(gNew, gamma, yNew, 1.0 + beta, -beta);
    // getResidual(gammaNew) (APGD.cu:90-102)
    P.schur(gNew, t);
    P.axpby(s.r.p, t, t, 1.0, 1.0);
    if (s.nc) {
      P.axpby(gNew, t, t, 1.0, -gdiff);
      P.project(t);
      P.axpby(gNew, t, t, 1.0 / gdiff, -1.0 / gdiff);
    }
    double res = P.nrmmax(t);
    if (!(res < INFINITY)) {  // non-finite data: the solve ends as diverged
      best = res;
      break;
    }
    if (res < best) {
      best = res;
      P.copy(gNew, gHat);
    }
    if (best < s.tol) break;
    P.axpby(gNew, gamma, t, 1.0, -1.0);
    if (P.dot(g, t) > 0) {
      P.copy(gNew, yNew);
      thetaNew = 1.0;
    }
    L = 0.9 * L;
    tt = 1.0 / L;
    theta = thetaNew;
    P.copy(gNew, gamma);
    P.copy(yNew, y);
    if (s.antiRelax) {
      // APGD.cu:335-345: r -= ar; f_contact = D^T gamma; v = M^-1 (k +
      // f_contact); ar from D v; r += ar.  Like the reference it leaves v and
      // f_contact overwritten (the update after the solve recomputes both)
      // and r holding the last anti-relaxation term.
      P.axpby(ar, s.r.p, s.r.p, -1.0, 1.0);
      gatherExact(s, gamma, s.fContact.p, nullptr, false);
      const long n3 = 3L * s.nb;
      if (n3) {
        hipLaunchKernelGGL(kAxpby, dim3(blocksFor(n3, kT)), dim3(kT), 0, s.stream, (const double*)s.k.p,
                           (const double*)s.fContact.p, s.tmp.p, 1.0, 1.0, n3);
        minvTimes(s, s.tmp.p, s.v.p);
      }
      rowsD(s, s.v.p, nullptr, t);
      if (s.nc) hipLaunchKernelGGL(kAntiRelax, dim3(blocksFor(s.nc, kT)), dim3(kT), 0, s.stream, (const double*)t, ar,
                                   s.nc, s.mu);
      P.axpby(ar, s.r.p, s.r.p, 1.0, 1.0);
    }
  }
  if (s.verbose) printf("  Iterations: %d Residual: %g\n", it, best);
  P.schur(gHat, t);
  s.stats.objective = 0.5 * P.dot(gHat, t) + P.dot(gHat, s.r.p);
  s.stats.iterations = it;
  s.stats.residual = best;
  s.stats.backtracks = backtracks;
  P.copy(gHat, gamma);
  DV_HIP(hipStreamSynchronize(s.stream));
}

}  // namespace dv
