// C ABI of libdvide_hip.so (include/dvide.h): scene lifetime, body upload
// (System::initializeSystem), the time step (System::DoTimeStep) and the
// data accessors.  Every entry point converts exceptions into DV_ERR_* codes
// and records the message for dv_last_error().
#include "dv_geom.h"
#include "dv_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

struct dv_scene {
  dv::Scene s;
};

namespace dv {
namespace {

template <class F>
int guarded(dv_scene* h, F&& fn) {
  if (!h) return DV_ERR_ARG;
  try {
    h->s.err.clear();
    return fn(h->s);
  } catch (const HipError& e) {
    h->s.err = e.where + ": " + hipGetErrorString(e.code);
    return DV_ERR_HIP;
  } catch (const DvError& e) {
    h->s.err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    h->s.err = e.what();
    return DV_ERR_STATE;
  }
}

int fail(Scene& s, int code, const char* msg) {
  s.err = msg;
  return code;
}

template <typename T>
void upload(DevBuf<T>& d, const std::vector<T>& h, hipStream_t st) {
  d.reserve(std::max<size_t>(h.size(), 1));
  if (!h.empty()) DV_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, st));
}

__global__ void kProjectN(double* g, long n, double mu) {
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double a = g[3 * c], b = g[3 * c + 1], d = g[3 * c + 2];
  projectCone<double>(a, b, d, mu);
  g[3 * c] = a; g[3 * c + 1] = b; g[3 * c + 2] = d;
}

void initialize(Scene& s) {
  if (s.initialized) throw DvError{DV_ERR_STATE, "initializeSystem called twice"};
  DV_HIP(hipGetDevice(&s.device));
  DV_HIP(hipDeviceGetAttribute(&s.numCU, hipDeviceAttributeMultiprocessorCount, s.device));
  if (!s.stream) DV_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
  for (auto& e : s.ev)
    if (!e) DV_HIP(hipEventCreate(&e));
  if (!s.pinned) DV_HIP(hipHostMalloc(&s.pinned, 16 * sizeof(int), hipHostMallocDefault));
  const int nb = (int)s.hMass.size();
  s.nb = nb;
  // System::initializeSystem (System.cu:344-424): f = m g (0 if fixed), M^-1 = 1/m (0 if fixed)
  std::vector<double> f(3 * (size_t)nb), minv(nb);
  for (int j = 0; j < nb; ++j) {
    for (int c = 0; c < 3; ++c) f[3 * j + c] = s.hFixed[j] ? 0.0 : s.hMass[j] * s.g[c];
    minv[j] = s.hFixed[j] ? 0.0 : 1.0 / s.hMass[j];
  }
  const hipStream_t st = s.stream;
  upload(s.p, s.hPos, st);
  upload(s.v, s.hVel, st);
  upload(s.f, f, st);
  upload(s.minv, minv, st);
  upload(s.geom, s.hGeom, st);
  upload(s.family, s.hFamily, st);
  s.k.reserve(3 * (size_t)std::max(nb, 1));
  s.fContact.reserve(3 * (size_t)std::max(nb, 1));
  s.tmp.reserve(3 * (size_t)std::max(nb, 1));
  fillDevice(s, s.fContact.p, 3L * nb, 0.0);
  // bilateral DOF constraints (System::processConstraints, System.cu:303-340)
  s.nbil = (int)s.hBilA.size();
  s.hBilP0.assign(s.nbil, 0.0);
  for (int i = 0; i < s.nbil; ++i) {
    if (s.hBilA[i] < 0 || s.hBilA[i] >= 3 * nb || s.hBilB[i] >= 3 * nb)
      throw DvError{DV_ERR_ARG, "bilateral constraint DOF out of range"};
    if (s.hBilB[i] < 0) s.hBilP0[i] = s.hPos[s.hBilA[i]];
  }
  upload(s.bilA, s.hBilA, st);
  upload(s.bilB, s.hBilB, st);
  upload(s.bilVel, s.hBilVel, st);
  upload(s.bilT0, s.hBilT0, st);
  upload(s.bilP0, s.hBilP0, st);
  // per-body bilateral incidences, rows ascending: (row << 3) | (comp << 1) | sign
  std::vector<std::vector<int>> per(nb);
  for (int i = 0; i < s.nbil; ++i) {
    per[s.hBilA[i] / 3].push_back((i << 3) | ((s.hBilA[i] % 3) << 1) | 0);
    if (s.hBilB[i] >= 0) per[s.hBilB[i] / 3].push_back((i << 3) | ((s.hBilB[i] % 3) << 1) | 1);
  }
  std::vector<int> bs(nb + 1, 0), bl;
  for (int j = 0; j < nb; ++j) {
    bs[j + 1] = bs[j] + (int)per[j].size();
    for (int e : per[j]) bl.push_back(e);
  }
  upload(s.bilIncStart, bs, st);
  upload(s.bilInc, bl, st);
  // gamma_d starts as zeros, one per DOF (System.cu:185)
  s.gamma.reserve(3 * (size_t)std::max(nb, 1));
  fillDevice(s, s.gamma.p, 3L * nb, 0.0);
  s.gammaLen = 3 * nb;
  collideInit(s);
  DV_HIP(hipStreamSynchronize(st));
  s.initialized = true;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
  return ms;
}

void runSolver(Scene& s) {
  switch (s.solver) {
    case DV_SOLVER_APGD_PLAIN: solvePlain(s); break;
    case DV_SOLVER_APGD_MIXED: solveFused(s, true); break;
    default: solveFused(s, false); break;
  }
  if (!std::isfinite(s.stats.residual) && s.stats.iterations > 0)
    throw DvError{DV_ERR_DIVERGED, "APGD produced a non-finite residual"};
}

void step(Scene& s) {
  if (!s.initialized) throw DvError{DV_ERR_STATE, "DoTimeStep before initializeSystem"};
  const hipStream_t st = s.stream;
  s.stats = dv_stats{};
  DV_HIP(hipEventRecord(s.ev[0], st));
  collide(s);
  DV_HIP(hipEventRecord(s.ev[1], st));
  buildK(s);
  const bool rows = (s.nc + s.nbil) > 0;
  if (rows) buildR(s);
  DV_HIP(hipEventRecord(s.ev[2], st));
  if (rows) runSolver(s);
  DV_HIP(hipEventRecord(s.ev[3], st));
  updateState(s, rows, s.solver == DV_SOLVER_APGD_PLAIN);
  DV_HIP(hipEventRecord(s.ev[4], st));
  DV_HIP(hipEventSynchronize(s.ev[4]));
  DV_HIP(hipGetLastError());
  s.time += s.h;
  s.timeIndex++;
  s.stats.num_contacts = s.nc;
  s.stats.num_bilaterals = s.nbil;
  s.stats.ms_collision = elapsed(s.ev[0], s.ev[1]);
  s.stats.ms_assembly = elapsed(s.ev[1], s.ev[2]);
  s.stats.ms_solve = elapsed(s.ev[2], s.ev[3]);
  s.stats.ms_update = elapsed(s.ev[3], s.ev[4]);
  s.stats.ms_total = elapsed(s.ev[0], s.ev[4]);
  s.stats.time = s.time;
  if (s.verbose)
    printf("Time: %f (Exec. Time: %f), Collisions: %d\n", s.time, s.stats.ms_total, s.nc);
}

}  // namespace
}  // namespace dv

using dv::Scene;

extern "C" {

namespace {
thread_local std::string g_createError;  // dv_last_error(NULL) after a failed create
}

int dv_scene_create(dv_scene** out, int solver) {
  if (!out) return DV_ERR_ARG;
  *out = nullptr;
  // no CPU fallback: without a HIP device there is nothing to run on
  int devices = 0;
  const hipError_t e = hipGetDeviceCount(&devices);
  if (e != hipSuccess || devices == 0) {
    g_createError = std::string("no HIP device visible (") + hipGetErrorString(e) + ")";
    return DV_ERR_HIP;
  }
  try {
    dv_scene* h = new dv_scene();
    const char* ng = getenv("DVIDE_NO_GRAPH");
    h->s.noGraph = ng && ng[0] == '1';
    h->s.solver = (solver == DV_SOLVER_APGD_MIXED || solver == DV_SOLVER_APGD_PLAIN) ? solver
                                                                                      : DV_SOLVER_APGD_FUSED;
    *out = h;
    return DV_OK;
  } catch (...) {
    return DV_ERR_NOMEM;
  }
}

int dv_scene_destroy(dv_scene* h) {
  if (!h) return DV_ERR_ARG;
  Scene& s = h->s;
  if (s.stream) (void)hipStreamSynchronize(s.stream);
  dv::DevBuf<double>* dbufs[] = {&s.p, &s.v, &s.f, &s.k, &s.fContact, &s.tmp, &s.minv, &s.geom, &s.bilVel,
                                 &s.bilT0, &s.bilP0, &s.cellPosR, &s.frame, &s.gap, &s.loc, &s.r,
                                 &s.b, &s.gamma, &s.work, &s.partials};
  for (auto* b : dbufs) b->release();
  dv::DevBuf<int>* ibufs[] = {&s.family, &s.smallIds, &s.bigIds, &s.bilA, &s.bilB, &s.bilIncStart,
                              &s.bilInc, &s.cellItems, &s.cntUp, &s.cntDown, &s.cStart, &s.iStart,
                              &s.scanTmp, &s.counters, &s.hostTotals, &s.cA, &s.cB, &s.inc, &s.heavy,
                              &s.bigChunks};
  for (auto* b : ibufs) b->release();
  s.slotOf.release(); s.slotCount.release(); s.slotStart.release(); s.slotFill.release();
  s.frame32.release(); s.work32.release();
  s.graphFp64.reset();
  s.graphMixed.reset();
  if (s.state) (void)hipFree(s.state);
  if (s.ticket) (void)hipFree(s.ticket);
  s.workM64.release();
  if (s.solveDims) (void)hipFree(s.solveDims);
  if (s.stateHost) (void)hipHostFree(s.stateHost);
  if (s.dimsHost) (void)hipHostFree(s.dimsHost);
  if (s.pinned) (void)hipHostFree(s.pinned);
  for (auto& e : s.ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : s.profEv) (void)hipEventDestroy(e);
  if (s.stream) (void)hipStreamDestroy(s.stream);
  delete h;
  return DV_OK;
}

const char* dv_last_error(const dv_scene* h) { return h ? h->s.err.c_str() : g_createError.c_str(); }

int dv_add_bodies(dv_scene* h, int n, const double* pos, const double* vel, const double* mass,
                  const double* geom, const int* fixed, const int* family) {
  return dv::guarded(h, [&](Scene& s) {
    if (s.initialized) return dv::fail(s, DV_ERR_STATE, "bodies must be added before initializeSystem");
    if (n < 0 || (n > 0 && (!pos || !mass || !geom))) return dv::fail(s, DV_ERR_ARG, "bad body arrays");
    for (int i = 0; i < n; ++i) {
      if (!(mass[i] > 0) && !(fixed && fixed[i]))
        return dv::fail(s, DV_ERR_ARG, "a free body needs a positive mass");
      if (!(geom[3 * i] > 0)) return dv::fail(s, DV_ERR_ARG, "geometry x extent must be positive");
    }
    for (int i = 0; i < n; ++i) {
      for (int c = 0; c < 3; ++c) {
        s.hPos.push_back(pos[3 * i + c]);
        s.hVel.push_back(vel ? vel[3 * i + c] : 0.0);
        s.hGeom.push_back(geom[3 * i + c]);
      }
      s.hMass.push_back(mass[i]);
      s.hFixed.push_back(fixed ? (fixed[i] ? 1 : 0) : 0);
      s.hFamily.push_back(family ? family[i] : -1);
    }
    return DV_OK;
  });
}

int dv_add_bilateral_dof(dv_scene* h, int dofA, int dofB, double velocity, double start_time) {
  return dv::guarded(h, [&](Scene& s) {
    if (s.initialized) return dv::fail(s, DV_ERR_STATE, "constraints must be added before initializeSystem");
    if (dofA < 0) return dv::fail(s, DV_ERR_ARG, "dofA must be >= 0");
    s.hBilA.push_back(dofA);
    s.hBilB.push_back(dofB < 0 ? -1 : dofB);
    s.hBilVel.push_back(velocity);
    s.hBilT0.push_back(start_time);
    return DV_OK;
  });
}

int dv_set_bins_per_axis(dv_scene* h, int bx, int by, int bz) {
  return dv::guarded(h, [&](Scene& s) {
    if (bx <= 0 || by <= 0 || bz <= 0) return dv::fail(s, DV_ERR_ARG, "bins must be positive");
    s.bins[0] = bx; s.bins[1] = by; s.bins[2] = bz;
    return DV_OK;
  });
}

int dv_set_param(dv_scene* h, int param, double v) {
  return dv::guarded(h, [&](Scene& s) {
    switch (param) {
      case DV_PARAM_H:
        if (!(v > 0)) return dv::fail(s, DV_ERR_ARG, "time step must be positive");
        s.h = v;
        break;
      case DV_PARAM_MU:
        if (!(v >= 0)) return dv::fail(s, DV_ERR_ARG, "friction must be >= 0");
        s.mu = v;
        break;
      case DV_PARAM_TOL: s.tol = v; break;
      case DV_PARAM_MAX_ITER:
        if (v < 0) return dv::fail(s, DV_ERR_ARG, "maxIterations must be >= 0");
        s.maxIt = (int)v;
        break;
      case DV_PARAM_GX:
      case DV_PARAM_GY:
      case DV_PARAM_GZ:
        if (s.initialized) return dv::fail(s, DV_ERR_STATE, "gravity is fixed at initializeSystem");
        s.g[param - DV_PARAM_GX] = v;
        break;
      case DV_PARAM_ENVELOPE:
        if (!(v >= 0)) return dv::fail(s, DV_ERR_ARG, "envelope must be >= 0");
        s.envelope = v;
        break;
      case DV_PARAM_WARM_START: s.warm = v != 0; break;
      case DV_PARAM_SOLVER:
        if ((int)v != DV_SOLVER_APGD_FUSED && (int)v != DV_SOLVER_APGD_MIXED && (int)v != DV_SOLVER_APGD_PLAIN)
          return dv::fail(s, DV_ERR_ARG, "unknown solver");
        s.solver = (int)v;
        break;
      case DV_PARAM_GEOM_MODE:
        if (v != 0 && v != 1) return dv::fail(s, DV_ERR_ARG, "geometry mode is 0 or 1");
        s.geomMode = (int)v;
        break;
      case DV_PARAM_TIME: s.time = v; break;
      case DV_PARAM_VERBOSE: s.verbose = (int)v; break;
      case DV_PARAM_GRID: s.grid = std::max(0, (int)v); break;
      case DV_PARAM_PROFILE:
        if (v < 0 || v > 4) return dv::fail(s, DV_ERR_ARG, "profile kernel is 0..4");
        s.profKernel = (int)v;
        break;
      default: return dv::fail(s, DV_ERR_ARG, "unknown parameter");
    }
    return DV_OK;
  });
}

int dv_get_param(const dv_scene* hc, int param, double* v) {
  dv_scene* h = const_cast<dv_scene*>(hc);
  return dv::guarded(h, [&](Scene& s) {
    if (!v) return DV_ERR_ARG;
    switch (param) {
      case DV_PARAM_H: *v = s.h; break;
      case DV_PARAM_MU: *v = s.mu; break;
      case DV_PARAM_TOL: *v = s.tol; break;
      case DV_PARAM_MAX_ITER: *v = s.maxIt; break;
      case DV_PARAM_GX: case DV_PARAM_GY: case DV_PARAM_GZ: *v = s.g[param - DV_PARAM_GX]; break;
      case DV_PARAM_ENVELOPE: *v = s.envelope; break;
      case DV_PARAM_WARM_START: *v = s.warm ? 1 : 0; break;
      case DV_PARAM_SOLVER: *v = s.solver; break;
      case DV_PARAM_GEOM_MODE: *v = s.geomMode; break;
      case DV_PARAM_TIME: *v = s.time; break;
      case DV_PARAM_VERBOSE: *v = s.verbose; break;
      case DV_PARAM_GRID: *v = s.grid; break;
      case DV_PARAM_PROFILE: *v = s.profKernel; break;
      default: return dv::fail(s, DV_ERR_ARG, "unknown parameter");
    }
    return DV_OK;
  });
}

int dv_initialize(dv_scene* h) {
  return dv::guarded(h, [&](Scene& s) {
    dv::initialize(s);
    return DV_OK;
  });
}

int dv_step(dv_scene* h) {
  return dv::guarded(h, [&](Scene& s) {
    dv::step(s);
    return DV_OK;
  });
}

int dv_num_bodies(const dv_scene* h, int* n) {
  if (!h || !n) return DV_ERR_ARG;
  *n = (int)h->s.hMass.size();
  return DV_OK;
}

int dv_num_contacts(const dv_scene* h, int* nc) {
  if (!h || !nc) return DV_ERR_ARG;
  *nc = h->s.nc;
  return DV_OK;
}

int dv_get_stats(const dv_scene* h, dv_stats* out) {
  if (!h || !out) return DV_ERR_ARG;
  *out = h->s.stats;
  return DV_OK;
}

int dv_get_vector(dv_scene* h, int which, double* out, size_t n, size_t* len) {
  return dv::guarded(h, [&](Scene& s) {
    if (!s.initialized) return dv::fail(s, DV_ERR_STATE, "not initialized");
    const double* src = nullptr;
    size_t L = 0;
    switch (which) {
      case DV_VEC_P: src = s.p.p; L = 3 * (size_t)s.nb; break;
      case DV_VEC_V: src = s.v.p; L = 3 * (size_t)s.nb; break;
      case DV_VEC_GAMMA: src = s.gamma.p; L = (size_t)s.gammaLen; break;
      case DV_VEC_F_CONTACT: src = s.fContact.p; L = 3 * (size_t)s.nb; break;
      case DV_VEC_R: src = s.r.p; L = (size_t)s.R; break;
      case DV_VEC_K: src = s.k.p; L = 3 * (size_t)s.nb; break;
      case DV_VEC_B: src = s.b.p; L = (size_t)s.R; break;
      default: return dv::fail(s, DV_ERR_ARG, "unknown vector");
    }
    if (len) *len = L;
    size_t m = std::min(n, L);
    if (out && m) {
      DV_HIP(hipMemcpyAsync(out, src, sizeof(double) * m, hipMemcpyDeviceToHost, s.stream));
      DV_HIP(hipStreamSynchronize(s.stream));
    }
    return DV_OK;
  });
}

int dv_set_vector(dv_scene* h, int which, const double* in, size_t n) {
  return dv::guarded(h, [&](Scene& s) {
    if (!s.initialized) return dv::fail(s, DV_ERR_STATE, "not initialized");
    double* dst = nullptr;
    size_t L = 3 * (size_t)s.nb;
    switch (which) {
      case DV_VEC_P: dst = s.p.p; break;
      case DV_VEC_V: dst = s.v.p; break;
      case DV_VEC_GAMMA: dst = s.gamma.p; L = (size_t)s.gammaLen; break;
      default: return dv::fail(s, DV_ERR_ARG, "only p, v and gamma can be written");
    }
    if (n != L || !in) return dv::fail(s, DV_ERR_ARG, "vector length mismatch");
    if (n) {
      DV_HIP(hipMemcpyAsync(dst, in, sizeof(double) * n, hipMemcpyHostToDevice, s.stream));
      DV_HIP(hipStreamSynchronize(s.stream));
    }
    return DV_OK;
  });
}

int dv_get_contacts(dv_scene* h, uint32_t* idA, uint32_t* idB, double* nap, double* loc, size_t n) {
  return dv::guarded(h, [&](Scene& s) {
    if (n < (size_t)s.nc) return dv::fail(s, DV_ERR_ARG, "buffers smaller than the contact count");
    const int nc = s.nc;
    if (nc == 0) return DV_OK;
    std::vector<double> fr(9 * (size_t)s.ncCap), gp(nc), lc(3 * (size_t)s.ncCap);
    const hipStream_t st = s.stream;
    if (idA) DV_HIP(hipMemcpyAsync(idA, s.cA.p, sizeof(int) * nc, hipMemcpyDeviceToHost, st));
    if (idB) DV_HIP(hipMemcpyAsync(idB, s.cB.p, sizeof(int) * nc, hipMemcpyDeviceToHost, st));
    DV_HIP(hipMemcpyAsync(fr.data(), s.frame.p, sizeof(double) * 9 * s.ncCap, hipMemcpyDeviceToHost, st));
    DV_HIP(hipMemcpyAsync(gp.data(), s.gap.p, sizeof(double) * nc, hipMemcpyDeviceToHost, st));
    if (loc) DV_HIP(hipMemcpyAsync(lc.data(), s.loc.p, sizeof(double) * 3 * s.ncCap, hipMemcpyDeviceToHost, st));
    DV_HIP(hipStreamSynchronize(st));
    for (int c = 0; c < nc; ++c) {
      if (nap) {
        nap[4 * c] = fr[c];
        nap[4 * c + 1] = fr[(size_t)s.ncCap + c];
        nap[4 * c + 2] = fr[2 * (size_t)s.ncCap + c];
        nap[4 * c + 3] = gp[c];
      }
      if (loc)
        for (int q = 0; q < 3; ++q) loc[3 * c + q] = lc[(size_t)q * s.ncCap + c];
    }
    return DV_OK;
  });
}

int dv_ccp_violation(dv_scene* h, double out[4]) {
  return dv::guarded(h, [&](Scene& s) {
    out[0] = out[1] = out[2] = out[3] = 0;
    if (s.nc == 0) return DV_OK;
    const size_t R = (size_t)s.R;
    dv::DevBuf<double> dvb;
    dvb.reserve(R);
    dv::rowsD(s, s.v.p, s.b.p, dvb.p);  // D v + b
    std::vector<double> w(R), g(R);
    DV_HIP(hipMemcpyAsync(w.data(), dvb.p, sizeof(double) * R, hipMemcpyDeviceToHost, s.stream));
    DV_HIP(hipMemcpyAsync(g.data(), s.gamma.p, sizeof(double) * R, hipMemcpyDeviceToHost, s.stream));
    DV_HIP(hipStreamSynchronize(s.stream));
    dvb.release();
    double mn = 1e300, mv = 1e300, comp = 0, cone = 0;
    for (int c = 0; c < s.nc; ++c) {
      const size_t f = (size_t)s.nbil + 3 * (size_t)c;
      const double gn = g[f], vn = w[f];
      mn = std::min(mn, gn);
      mv = std::min(mv, vn);
      comp += gn * vn;
      double gt = std::sqrt(g[f + 1] * g[f + 1] + g[f + 2] * g[f + 2]);
      double cv = s.mu * gn - gt;
      if (cv > 0) cv = 0;
      cone += cv * cv;
    }
    out[0] = std::min(mn, 0.0);
    out[1] = std::min(mv, 0.0);
    out[2] = comp;
    out[3] = std::sqrt(cone);
    return DV_OK;
  });
}

int dv_assemble(dv_scene* h, int* nc) {
  return dv::guarded(h, [&](Scene& s) {
    if (!s.initialized) return dv::fail(s, DV_ERR_STATE, "not initialized");
    dv::collide(s);
    dv::buildK(s);
    if (s.nc + s.nbil) dv::buildR(s);
    DV_HIP(hipStreamSynchronize(s.stream));
    if (nc) *nc = s.nc;
    return DV_OK;
  });
}

int dv_schur_product(dv_scene* h, const double* x, double* y, size_t n) {
  return dv::guarded(h, [&](Scene& s) {
    if (n != (size_t)s.R || !x || !y) return dv::fail(s, DV_ERR_ARG, "x and y need nbil + 3*nc entries");
    if (n == 0) return DV_OK;
    dv::DevBuf<double> buf;
    buf.reserve(2 * n);
    DV_HIP(hipMemcpyAsync(buf.p, x, sizeof(double) * n, hipMemcpyHostToDevice, s.stream));
    dv::schurExact(s, buf.p, buf.p + n);
    DV_HIP(hipMemcpyAsync(y, buf.p + n, sizeof(double) * n, hipMemcpyDeviceToHost, s.stream));
    DV_HIP(hipStreamSynchronize(s.stream));
    buf.release();
    return DV_OK;
  });
}

int dv_project_cone(dv_scene* h, double* g, size_t n, double mu) {
  return dv::guarded(h, [&](Scene& s) {
    if (!g) return DV_ERR_ARG;
    if (n == 0) return DV_OK;
    if (!s.stream) DV_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    dv::DevBuf<double> buf;
    buf.reserve(3 * n);
    DV_HIP(hipMemcpyAsync(buf.p, g, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s.stream));
    hipLaunchKernelGGL(dv::kProjectN, dim3(dv::blocksFor((long)n, 256)), dim3(256), 0, s.stream, buf.p, (long)n, mu);
    DV_HIP(hipMemcpyAsync(g, buf.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s.stream));
    DV_HIP(hipStreamSynchronize(s.stream));
    buf.release();
    return DV_OK;
  });
}

int dv_solve(dv_scene* h) {
  return dv::guarded(h, [&](Scene& s) {
    if (!s.initialized) return dv::fail(s, DV_ERR_STATE, "not initialized");
    if (s.nc + s.nbil == 0) return DV_OK;
    dv::runSolver(s);
    return DV_OK;
  });
}

}  // extern "C"
