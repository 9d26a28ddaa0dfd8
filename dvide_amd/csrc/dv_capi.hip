// C ABI of libdvide_hip.so (include/dvide.h): scene lifetime, body upload
// (System::initializeSystem), the time step (System::DoTimeStep) and the
// data accessors.  Every entry point converts exceptions into DV_ERR_* codes
// and records the message for dv_last_error().
#include "dv_geom.h"
#include "dv_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <exception>
#include <fstream>
#include <mutex>
#include <new>
#include <sstream>
#include <thread>

struct dv_scene {
  dv::Scene s;
};

// Staging of dv_schur_product_noalloc: x in the first kNoAllocRows doubles,
// y in the second.  Every GPU the library runs on holds its own copy (16 MiB)
// until the process ends.  In host code `scratch` names a host-side stand-in:
// the device address of each GPU's copy is taken once per device, at scene
// initialisation, and kept in scratchPtr.
static constexpr size_t kNoAllocRows = size_t(1) << 20;
static __device__ __align__(256) double scratch[2 * kNoAllocRows];

namespace dv {
namespace {

constexpr int kMaxDevices = 64;
std::mutex scratchMutex;
double* scratchPtr[kMaxDevices] = {};

void resolveScratch(int device) {
  if (device < 0 || device >= kMaxDevices) throw DvError{DV_ERR_ARG, "device index beyond the scratch table"};
  std::lock_guard<std::mutex> lock(scratchMutex);
  if (scratchPtr[device]) return;
  void* p = nullptr;
  DV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));  // the current device's copy
  scratchPtr[device] = static_cast<double*>(p);
}

double* scratchOf(int device) {
  std::lock_guard<std::mutex> lock(scratchMutex);
  return (device >= 0 && device < kMaxDevices) ? scratchPtr[device] : nullptr;
}

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
  resolveScratch(s.device);  // with s.device current
  if (!s.stream) DV_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
  for (auto& e : s.ev)
    if (!e) DV_HIP(hipEventCreate(&e));
  if (!s.pinned) DV_HIP(hipHostMalloc(&s.pinned, 16 * sizeof(int), hipHostMallocDefault));
  const long side = s.nNodes ? s.shellPoints - 2 : 0, nPts = (long)(s.hShellConn.size() / 4) * side * side;
  if ((long)s.hMass.size() + nPts > kMaxBodies)
    throw DvError{DV_ERR_UNSUPPORTED, "more than 2^28 bodies and shell contact points in one scene"};
  s.nUser = (int)s.hMass.size();
  if (s.nNodes) shellAppendPoints(s);  // the shells' contact points follow the caller's bodies
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
  if (s.nNodes) shellPlacePoints(s);
  // bilateral DOF constraints (System::processConstraints, System.cu:303-340)
  s.nbil = (int)s.hBilA.size();
  s.hBilP0.assign(s.nbil, 0.0);
  for (int i = 0; i < s.nbil; ++i) {
    const int ndof = 3 * s.nUser + 9 * s.nNodes;
    if (s.hBilA[i] < 0 || s.hBilA[i] >= ndof || s.hBilB[i] >= ndof)
      throw DvError{DV_ERR_ARG, "bilateral constraint DOF out of range"};
    if (s.hBilA[i] >= 3 * s.nUser || s.hBilB[i] >= 3 * s.nUser)
      throw DvError{DV_ERR_UNSUPPORTED, "bilateral constraints on shell DOFs (the shells are fixed)"};
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
  // gamma_d starts as zeros, one per DOF (System.cu:185), the shells' included
  const long ndof = 3L * s.nUser + 9L * s.nNodes;
  s.gamma.reserve((size_t)std::max(ndof, 1L));
  fillDevice(s, s.gamma.p, ndof, 0.0);
  s.gammaLen = (int)ndof;
  collideInit(s);
  DV_HIP(hipStreamSynchronize(st));
  s.initialized = true;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
  return ms;
}

// The build that solves the scene: anti-relaxation runs on the plain build,
// which a DV_APGD_FP64 scene then uses (the mixed build rejects it).
int solveBuild(const Scene& s) {
  if (s.antiRelax && s.build == DV_APGD_MIXED)
    throw DvError{DV_ERR_UNSUPPORTED, "anti-relaxation runs on the FP64 builds, not the mixed one"};
  return s.antiRelax ? DV_APGD_PLAIN : s.build;
}

void runSolver(Scene& s) {
  switch (solveBuild(s)) {
    case DV_APGD_PLAIN: solvePlain(s); break;
    case DV_APGD_MIXED: solveFused(s, true); break;
    default: solveFused(s, false); break;
  }
  if (!std::isfinite(s.stats.residual))
    throw DvError{DV_ERR_DIVERGED, "APGD produced a non-finite residual"};
}

// First half of DoTimeStep (System.cu:544-568): collision detection and the
// assembly of k, D and r on the scene's stream; returns whether there are
// rows to solve.  stepAssembleBegin only queues the collision counts, so that
// a batch can queue those of all its scenes before waiting on any.
void stepAssembleBegin(Scene& s) {
  if (!s.initialized) throw DvError{DV_ERR_STATE, "DoTimeStep before initializeSystem"};
  s.stats = dv_stats{};
  DV_HIP(hipEventRecord(s.ev[0], s.stream));
  collideCount(s);
}

bool stepAssembleEnd(Scene& s) {
  const hipStream_t st = s.stream;
  collideEmit(s);
  DV_HIP(hipEventRecord(s.ev[1], st));
  buildK(s);
  const bool rows = (s.nc + s.nbil) > 0;
  if (rows) buildR(s);
  DV_HIP(hipEventRecord(s.ev[2], st));
  return rows;
}

bool stepAssemble(Scene& s) {
  stepAssembleBegin(s);
  return stepAssembleEnd(s);
}

// Second half (System.cu:569-590): the update after the solve, time and
// statistics.  The solve's results must be ordered before this on the
// scene's stream.  stepFinishQueue only queues the update, so that a batch
// can queue those of all its scenes before waiting on any.
void stepFinishQueue(Scene& s, bool rows, bool exactOrder) {
  if (rows && !std::isfinite(s.stats.residual))
    throw DvError{DV_ERR_DIVERGED, "APGD produced a non-finite residual"};
  const hipStream_t st = s.stream;
  DV_HIP(hipEventRecord(s.ev[3], st));
  updateState(s, rows, exactOrder);
  DV_HIP(hipEventRecord(s.ev[4], st));
}

void stepFinishWait(Scene& s) {
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

void stepFinish(Scene& s, bool rows, bool exactOrder) {
  stepFinishQueue(s, rows, exactOrder);
  stepFinishWait(s);
}

void step(Scene& s) {
  const bool rows = stepAssemble(s);
  if (rows) runSolver(s);
  stepFinish(s, rows, solveBuild(s) == DV_APGD_PLAIN);
}

}  // namespace
}  // namespace dv

using dv::Scene;

extern "C" {

namespace {
thread_local std::string g_createError;  // dv_last_error(NULL) after a failed create
}

int dv_scene_create(dv_scene** out, int solverType) {
  if (!out) return DV_ERR_ARG;
  *out = nullptr;
  // System.cu:77-101: 2..7 are PDIP, TPAS, JKIP, PGJ, PGS and PJKIP
  if (solverType >= 2 && solverType <= 7) {
    static const char* names[] = {"PDIP", "TPAS", "JKIP", "PGJ", "PGS", "PJKIP"};
    g_createError = std::string("solver type ") + std::to_string(solverType) + " (" + names[solverType - 2] +
                    ") is not on this path; only APGD (1) is";
    return DV_ERR_UNSUPPORTED;
  }
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
                                 &s.b, &s.gamma, &s.work, &s.partials, &s.shellNodes, &s.ptS, &s.shellF};
  for (auto* b : dbufs) b->release();
  dv::DevBuf<int>* ibufs[] = {&s.family, &s.smallIds, &s.bigIds, &s.bilA, &s.bilB, &s.bilIncStart,
                              &s.bilInc, &s.cellItems, &s.cntUp, &s.cntDown, &s.cStart, &s.iStart,
                              &s.scanTmp, &s.counters, &s.hostTotals, &s.cA, &s.cB, &s.inc, &s.posB,
                              &s.heavy, &s.heavyPos, &s.bigChunks, &s.shellConn, &s.ptEl, &s.nodeIncStart,
                              &s.nodeInc};
  for (auto* b : ibufs) b->release();
  s.slotOf.release(); s.slotCount.release(); s.slotStart.release(); s.slotFill.release();
  s.frame32.release(); s.work32.release();
  s.engine.release();
  s.workM64.release();
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

int dv_add_shell_mesh(dv_scene* h, int n_nodes, const double* nodes, int n_elements, const int* connectivity,
                      const double* geometry, int points_per_side, const double* fext) {
  return dv::guarded(h, [&](Scene& s) {
    if (s.initialized) return dv::fail(s, DV_ERR_STATE, "shell meshes must be added before initializeSystem");
    if (s.nNodes) return dv::fail(s, DV_ERR_UNSUPPORTED, "one shell mesh per scene");
    if (n_nodes <= 0 || n_elements <= 0 || !nodes || !connectivity || !geometry)
      return dv::fail(s, DV_ERR_ARG, "bad shell mesh arrays");
    if (points_per_side < 2) return dv::fail(s, DV_ERR_ARG, "points per element side must be at least 2");
    // the points join the bodies, and a node's incidence list packs (point, corner) into an int
    const long side = (long)points_per_side - 2;
    if (side > 8192 || side * side * (long)n_elements > dv::kMaxShellPoints)
      return dv::fail(s, DV_ERR_ARG, "more than 2^26 shell contact points");
    for (int e = 0; e < n_elements; ++e) {
      for (int q = 0; q < 4; ++q)
        if (connectivity[4 * e + q] < 0 || connectivity[4 * e + q] >= n_nodes)
          return dv::fail(s, DV_ERR_ARG, "shell connectivity names a node out of range");
      for (int q = 0; q < 3; ++q)
        if (!(geometry[4 * e + q] > 0)) return dv::fail(s, DV_ERR_ARG, "shell sides and thickness must be positive");
    }
    for (long i = 0; i < 9L * n_nodes; ++i)
      if (!std::isfinite(nodes[i]) || (fext && !std::isfinite(fext[i])))
        return dv::fail(s, DV_ERR_ARG, "non-finite shell node data");
    s.nNodes = n_nodes;
    s.hShellNodes.assign(nodes, nodes + 9L * n_nodes);
    if (fext) s.hShellFext.assign(fext, fext + 9L * n_nodes);
    else s.hShellFext.assign(9L * n_nodes, 0.0);
    s.hShellConn.assign(connectivity, connectivity + 4L * n_elements);
    s.hShellGeom.clear();
    for (int e = 0; e < n_elements; ++e)
      for (int q = 0; q < 3; ++q) s.hShellGeom.push_back(geometry[4 * e + q]);
    s.shellPoints = points_per_side;
    return DV_OK;
  });
}

// System::importMesh's reader (System.cu:1932-2060): one record per line,
// fields separated by commas or blanks; the arrays it fills go to
// dv_add_shell_mesh
static int importMesh(dv_scene* h, Scene& s, const char* path, int points_per_side) {
  if (!path) return dv::fail(s, DV_ERR_ARG, "null path");
  std::ifstream in(path);
  if (!in) return dv::fail(s, DV_ERR_ARG, (std::string("cannot open ") + path).c_str());
  // one record per line, fields separated by commas or blanks
  std::string line;
  auto next = [&](std::istringstream& ss) {
    if (!std::getline(in, line)) return false;
    std::replace(line.begin(), line.end(), ',', ' ');
    ss.clear();
    ss.str(line);
    return true;
  };
  std::istringstream ss;
  long nn = 0, ne = 0, nM = 0, nInv = 0;
  if (!next(ss) || !(ss >> nn >> ne >> nM >> nInv) || nn <= 0 || ne <= 0 || nM < 0 || nInv < 0)
    return dv::fail(s, DV_ERR_ARG, "bad mesh header (nodes, shells, M entries, M^-1 entries)");
  if (nn > (1L << 24) || ne > (1L << 24)) return dv::fail(s, DV_ERR_ARG, "mesh larger than 2^24 nodes or shells");
  std::vector<double> nodes(9 * nn), fext(9 * nn), geom(4 * ne);
  std::vector<int> conn(4 * ne);
  for (long i = 0; i < 3 * nn; ++i)
    if (!next(ss) || !(ss >> nodes[3 * i] >> nodes[3 * i + 1] >> nodes[3 * i + 2]))
      return dv::fail(s, DV_ERR_ARG, "bad mesh node line");
  for (long e = 0; e < ne; ++e)
    if (!next(ss) || !(ss >> conn[4 * e] >> conn[4 * e + 1] >> conn[4 * e + 2] >> conn[4 * e + 3]))
      return dv::fail(s, DV_ERR_ARG, "bad mesh connectivity line");
  for (long e = 0; e < ne; ++e) {  // materials (used by the elastic forces only)
    double m[4];
    if (!next(ss) || !(ss >> m[0] >> m[1] >> m[2] >> m[3])) return dv::fail(s, DV_ERR_ARG, "bad mesh material line");
  }
  for (long e = 0; e < ne; ++e)
    if (!next(ss) || !(ss >> geom[4 * e] >> geom[4 * e + 1] >> geom[4 * e + 2] >> geom[4 * e + 3]))
      return dv::fail(s, DV_ERR_ARG, "bad mesh geometry line");
  for (long i = 0; i < 36 * ne; ++i) {  // the element-to-DOF map (used by the elastic forces only)
    long m;
    if (!next(ss) || !(ss >> m)) return dv::fail(s, DV_ERR_ARG, "bad mesh map line");
  }
  for (long i = 0; i < 9 * nn; ++i)
    if (!next(ss) || !(ss >> fext[i])) return dv::fail(s, DV_ERR_ARG, "bad mesh external force line");
  for (long i = 0; i < nM; ++i) {  // M: the shell's momentum, which a fixed shell does not have
    long r, c;
    double v;
    if (!next(ss) || !(ss >> r >> c >> v)) return dv::fail(s, DV_ERR_ARG, "bad mesh mass line");
  }
  for (long i = 0; i < nInv; ++i) {
    long r, c;
    double v;
    if (!next(ss) || !(ss >> r >> c >> v)) return dv::fail(s, DV_ERR_ARG, "bad mesh inverse-mass line");
    if (v != 0)
      return dv::fail(s, DV_ERR_UNSUPPORTED,
                      "the mesh's inverse mass is not zero: free shells need their elastic forces, which this "
                      "library does not compute (only fixed shells are supported)");
  }
  return dv_add_shell_mesh(h, (int)nn, nodes.data(), (int)ne, conn.data(), geom.data(), points_per_side,
                           fext.data());
}

int dv_import_mesh(dv_scene* h, const char* path, double stiffness, int points_per_side, double pressure) {
  (void)stiffness;  // the elastic and pressure forces act on free shell DOFs only
  (void)pressure;
  return dv::guarded(h, [&](Scene& s) { return importMesh(h, s, path, points_per_side); });
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
        s.envelope = v;  // the hash cell edge 2 (r_max + envelope) follows it each step
        break;
      case DV_PARAM_WARM_START: s.warm = v != 0; break;
      case DV_PARAM_APGD_BUILD:
        if ((int)v != DV_APGD_FP64 && (int)v != DV_APGD_MIXED && (int)v != DV_APGD_PLAIN)
          return dv::fail(s, DV_ERR_ARG, "unknown APGD build");
        s.build = (int)v;
        break;
      case DV_PARAM_ANTI_RELAX: s.antiRelax = v != 0; break;
      case DV_PARAM_GEOM_MODE:
        if (v != 0 && v != 1) return dv::fail(s, DV_ERR_ARG, "geometry mode is 0 or 1");
        s.geomMode = (int)v;
        break;
      case DV_PARAM_TIME: s.time = v; break;
      case DV_PARAM_VERBOSE: s.verbose = (int)v; break;
      case DV_PARAM_GRID: s.grid = std::max(0, (int)v); break;
      case DV_PARAM_PROFILE:
        if (v < 0 || v > 4 || v == 3) return dv::fail(s, DV_ERR_ARG, "profile kernel is 0, 1, 2 or 4");
        s.profKernel = (int)v;
        break;
      case DV_PARAM_ROWS_PER_THREAD:
        if (v != 0 && v != 1 && v != 2) return dv::fail(s, DV_ERR_ARG, "rows per thread is 0 (tuned), 1 or 2");
        s.rowsPerThread = (int)v;
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
      case DV_PARAM_APGD_BUILD: *v = s.build; break;
      case DV_PARAM_ANTI_RELAX: *v = s.antiRelax ? 1 : 0; break;
      case DV_PARAM_GEOM_MODE: *v = s.geomMode; break;
      case DV_PARAM_TIME: *v = s.time; break;
      case DV_PARAM_VERBOSE: *v = s.verbose; break;
      case DV_PARAM_GRID: *v = s.grid; break;
      case DV_PARAM_PROFILE: *v = s.profKernel; break;
      case DV_PARAM_ROWS_PER_THREAD: *v = s.rowsPerThread; break;
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

// the caller's bodies (after initialisation the host arrays also hold the
// shells' contact points)
static int userBodies(const Scene& s) { return s.initialized ? s.nUser : (int)s.hMass.size(); }

int dv_num_bodies(const dv_scene* h, int* n) {
  if (!h || !n) return DV_ERR_ARG;
  *n = userBodies(h->s);
  return DV_OK;
}

int dv_num_shell_nodes(const dv_scene* h, int* n) {
  if (!h || !n) return DV_ERR_ARG;
  *n = h->s.nNodes;
  return DV_OK;
}

int dv_get_bodies(const dv_scene* h, double* mass, double* geom, int* fixed, int* family, size_t n) {
  if (!h) return DV_ERR_ARG;
  const Scene& s = h->s;
  const size_t m = std::min(n, (size_t)userBodies(s));
  for (size_t i = 0; i < m; ++i) {
    if (mass) mass[i] = s.hMass[i];
    if (fixed) fixed[i] = s.hFixed[i];
    if (family) family[i] = s.hFamily[i];
    if (geom)
      for (int q = 0; q < 3; ++q) geom[3 * i + q] = s.hGeom[3 * i + q];
  }
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
    // DOF vectors: the caller's bodies, then the shells' 9 DOFs per node
    const size_t nBody = 3 * (size_t)s.nUser, nShell = 9 * (size_t)s.nNodes;
    bool dof = true;
    switch (which) {
      case DV_VEC_P: src = s.p.p; break;
      case DV_VEC_V: src = s.v.p; break;
      case DV_VEC_F_CONTACT: src = s.fContact.p; break;
      case DV_VEC_K:
        if (nShell) return dv::fail(s, DV_ERR_UNSUPPORTED, "k of a scene with shells (their elastic forces are not computed)");
        src = s.k.p;
        break;
      case DV_VEC_GAMMA: src = s.gamma.p; L = (size_t)s.gammaLen; dof = false; break;
      case DV_VEC_R: src = s.r.p; L = (size_t)s.R; dof = false; break;
      case DV_VEC_B: src = s.b.p; L = (size_t)s.R; dof = false; break;
      default: return dv::fail(s, DV_ERR_ARG, "unknown vector");
    }
    if (dof) L = nBody + nShell;
    if (len) *len = L;
    if (!out || !n) return DV_OK;
    const size_t m = std::min(n, dof ? nBody : L);
    if (m) DV_HIP(hipMemcpyAsync(out, src, sizeof(double) * m, hipMemcpyDeviceToHost, s.stream));
    std::vector<double> sh;
    if (dof && n > nBody && nShell) {  // the shell DOFs: fixed nodes, no velocity
      if (which == DV_VEC_P) sh = s.hShellNodes;
      else if (which == DV_VEC_V) sh.assign(nShell, 0.0);
      else {
        sh.resize(nShell);
        dv::shellForces(s, s.shellF.p);
        DV_HIP(hipMemcpyAsync(sh.data(), s.shellF.p, sizeof(double) * nShell, hipMemcpyDeviceToHost, s.stream));
      }
    }
    DV_HIP(hipStreamSynchronize(s.stream));
    if (!sh.empty()) std::copy(sh.begin(), sh.begin() + std::min(nShell, n - nBody), out + nBody);
    return DV_OK;
  });
}

int dv_set_vector(dv_scene* h, int which, const double* in, size_t n) {
  return dv::guarded(h, [&](Scene& s) {
    if (!s.initialized) return dv::fail(s, DV_ERR_STATE, "not initialized");
    double* dst = nullptr;
    const size_t nBody = 3 * (size_t)s.nUser, nShell = 9 * (size_t)s.nNodes;
    size_t L = nBody + nShell;
    switch (which) {
      case DV_VEC_P: dst = s.p.p; break;
      case DV_VEC_V: dst = s.v.p; break;
      case DV_VEC_GAMMA: dst = s.gamma.p; L = (size_t)s.gammaLen; break;
      default: return dv::fail(s, DV_ERR_ARG, "only p, v and gamma can be written");
    }
    if (n != L || !in) return dv::fail(s, DV_ERR_ARG, "vector length mismatch");
    if (which != DV_VEC_GAMMA && nShell) {  // the shells are fixed: their DOFs must keep their values
      for (size_t i = 0; i < nShell; ++i)
        if (in[nBody + i] != (which == DV_VEC_P ? s.hShellNodes[i] : 0.0))
          return dv::fail(s, DV_ERR_UNSUPPORTED, "the shell DOFs are fixed");
      n = nBody;
    }
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

int dv_get_jacobian(dv_scene* h, int* I, int* J, double* V, size_t cap, size_t* nnz, int* rows, int* cols) {
  return dv::guarded(h, [&](Scene& s) {
    if (!s.initialized) return dv::fail(s, DV_ERR_STATE, "not initialized");
    const int nc = s.nc, nbil = (int)s.hBilA.size();
    size_t bilNnz = 0;
    for (int q = 0; q < nbil; ++q) bilNnz += s.hBilB[q] >= 0 ? 2 : 1;
    std::vector<double> fr(9 * (size_t)std::max(s.ncCap, 1));
    std::vector<int> a(std::max(nc, 1)), b(std::max(nc, 1));
    if (nc) {
      DV_HIP(hipMemcpyAsync(fr.data(), s.frame.p, sizeof(double) * 9 * s.ncCap, hipMemcpyDeviceToHost, s.stream));
      DV_HIP(hipMemcpyAsync(a.data(), s.cA.p, sizeof(int) * nc, hipMemcpyDeviceToHost, s.stream));
      DV_HIP(hipMemcpyAsync(b.data(), s.cB.p, sizeof(int) * nc, hipMemcpyDeviceToHost, s.stream));
      DV_HIP(hipStreamSynchronize(s.stream));
    }
    // a side of a contact row: a body's 3 DOFs, or the 36 DOFs of a shell
    // point's element, node by node (position, x-gradient, y-gradient, each
    // x, y, z) weighted by the shape functions (System.cu:785-800, 904-961)
    const int nU = s.nUser;
    auto width = [&](int g) { return g < nU ? 3 : 36; };
    size_t total = bilNnz;
    for (int c = 0; c < nc; ++c) total += 3 * (size_t)(width(a[c]) + width(b[c]));
    if (nnz) *nnz = total;
    if (rows) *rows = 3 * nc + nbil;
    if (cols) *cols = 3 * nU + 9 * s.nNodes;
    if (!I || !J || !V) return DV_OK;
    if (cap < total) return dv::fail(s, DV_ERR_ARG, "buffers smaller than nnz");
    size_t o = 0;
    for (int q = 0; q < nbil; ++q) {  // System.cu:622-638
      I[o] = q; J[o] = s.hBilA[q]; V[o] = 1.0; ++o;
      if (s.hBilB[q] >= 0) { I[o] = q; J[o] = s.hBilB[q]; V[o] = -1.0; ++o; }
    }
    auto side = [&](int g, int row, const double* val, double sign) {
      if (g < nU) {
        for (int q = 0; q < 3; ++q) { I[o] = row; J[o] = 3 * g + q; V[o] = sign * val[q]; ++o; }
        return;
      }
      const int j = g - nU, e = s.hPtEl[j];
      const double* S = &s.hPtS[12 * (size_t)j];
      for (int q = 0; q < 4; ++q)
        for (int m = 0; m < 9; ++m) {
          I[o] = row;
          J[o] = 3 * nU + 9 * s.hShellConn[4 * e + q] + m;
          V[o] = sign * (S[3 * q + m / 3] * val[m % 3]);
          ++o;
        }
    };
    for (int c = 0; c < nc; ++c) {
      for (int rr = 0; rr < 3; ++rr) {
        const int row = nbil + 3 * c + rr;
        double val[3];
        for (int q = 0; q < 3; ++q) val[q] = fr[(size_t)(3 * rr + q) * s.ncCap + c];
        side(a[c], row, val, 1.0);
        side(b[c], row, val, -1.0);
      }
    }
    return DV_OK;
  });
}

int dv_contact_forces(dv_scene* h, double* normal, double* friction, size_t n) {
  return dv::guarded(h, [&](Scene& s) {
    if (!normal || !friction) return dv::fail(s, DV_ERR_ARG, "null buffer");
    if (n < (size_t)s.nc) return dv::fail(s, DV_ERR_ARG, "buffers smaller than the contact count");
    const int nc = s.nc;
    if (nc == 0) return DV_OK;
    std::vector<double> fr(9 * (size_t)s.ncCap), g((size_t)s.gammaLen);
    DV_HIP(hipMemcpyAsync(fr.data(), s.frame.p, sizeof(double) * 9 * s.ncCap, hipMemcpyDeviceToHost, s.stream));
    DV_HIP(hipMemcpyAsync(g.data(), s.gamma.p, sizeof(double) * g.size(), hipMemcpyDeviceToHost, s.stream));
    DV_HIP(hipStreamSynchronize(s.stream));
    const double h = s.h;
    for (int c = 0; c < nc; ++c) {
      double F[9];
      for (int k = 0; k < 9; ++k) F[k] = fr[(size_t)k * s.ncCap + c];
      const size_t i0 = 3 * (size_t)c;  // no bilateral offset (System.cu:1279)
      const double g0 = i0 < g.size() ? g[i0] : 0, g1 = i0 + 1 < g.size() ? g[i0 + 1] : 0,
                   g2 = i0 + 2 < g.size() ? g[i0 + 2] : 0;
      for (int q = 0; q < 3; ++q) {
        normal[3 * c + q] = g0 * F[q] / h;
        friction[3 * c + q] = (g1 * F[3 + q] + g2 * F[6 + q]) / h;
      }
    }
    return DV_OK;
  });
}

int dv_energies(dv_scene* h, double out[3]) {
  return dv::guarded(h, [&](Scene& s) {
    if (!s.initialized) return dv::fail(s, DV_ERR_STATE, "not initialized");
    const size_t n3 = 3 * (size_t)s.nb;
    std::vector<double> p(n3), v(n3), f(n3), mi(s.nb);
    DV_HIP(hipMemcpyAsync(p.data(), s.p.p, sizeof(double) * n3, hipMemcpyDeviceToHost, s.stream));
    DV_HIP(hipMemcpyAsync(v.data(), s.v.p, sizeof(double) * n3, hipMemcpyDeviceToHost, s.stream));
    DV_HIP(hipMemcpyAsync(f.data(), s.f.p, sizeof(double) * n3, hipMemcpyDeviceToHost, s.stream));
    DV_HIP(hipMemcpyAsync(mi.data(), s.minv.p, sizeof(double) * s.nb, hipMemcpyDeviceToHost, s.stream));
    DV_HIP(hipStreamSynchronize(s.stream));
    double pe = 0, ke = 0;
    for (size_t i = 0; i < 3 * (size_t)s.nUser; ++i) {
      pe += f[i] * p[i];
      const double m = mi[i / 3] != 0 ? 1.0 / mi[i / 3] : 0.0;  // multiplyByMass, System.cu:1411-1417
      ke += v[i] * (m * v[i]);
    }
    // the fixed shells: f is their external force, v zero (no kinetic energy)
    for (size_t i = 0; i < s.hShellFext.size(); ++i) pe += s.hShellFext[i] * s.hShellNodes[i];
    out[0] = -pe;
    out[1] = 0.5 * ke;
    out[2] = 0.0;
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

int dv_pass_product(dv_scene* h, int build, const double* x, double* y, size_t n) {
  return dv::guarded(h, [&](Scene& s) {
    if (build != DV_APGD_FP64 && build != DV_APGD_MIXED)
      return dv::fail(s, DV_ERR_ARG, "the pass kernels belong to the FP64 and mixed builds");
    if (!s.initialized) return dv::fail(s, DV_ERR_STATE, "not initialized");
    if (n != (size_t)s.R || !x || !y) return dv::fail(s, DV_ERR_ARG, "x and y need nbil + 3*nc entries");
    if (n == 0) return DV_OK;
    dv::DevBuf<double> xs;
    xs.reserve(n);
    DV_HIP(hipMemcpyAsync(xs.p, x, sizeof(double) * n, hipMemcpyHostToDevice, s.stream));
    dv::passProduct(s, build == DV_APGD_MIXED, xs.p, y);
    xs.release();
    return DV_OK;
  });
}

int dv_schur_product_noalloc(dv_scene* h, const double* x, double* y, size_t n) {
  return dv::guarded(h, [&](Scene& s) {
    if (!x || !y) return dv::fail(s, DV_ERR_ARG, "null x or y");
    if (n > kNoAllocRows)
      return dv::fail(s, DV_ERR_ARG, "x and y exceed the 1,048,576 rows of the no-allocation scratch");
    if (!s.initialized) return dv::fail(s, DV_ERR_STATE, "not initialized");
    if (n != (size_t)s.R) return dv::fail(s, DV_ERR_ARG, "x and y need nbil + 3*nc entries");
    if (n == 0) return DV_OK;
    double* base = dv::scratchOf(s.device);  // this device's copy of `scratch`
    if (!base) return dv::fail(s, DV_ERR_STATE, "scratch of the scene's device not resolved");
    // x is overwritten in full and every row of y is written, so nothing in
    // the scratch needs clearing
    DV_HIP(hipMemcpyAsync(base, x, sizeof(double) * n, hipMemcpyHostToDevice, s.stream));
    dv::schurExact(s, base, base + kNoAllocRows);
    DV_HIP(hipMemcpyAsync(y, base + kNoAllocRows, sizeof(double) * n, hipMemcpyDeviceToHost, s.stream));
    DV_HIP(hipStreamSynchronize(s.stream));
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

// ---------------------------------------------------------------------------
// batches and serving

struct dv_batch {
  int build = DV_APGD_MIXED;
  std::vector<dv_scene*> scenes;
  dv::Engine engine;
  hipStream_t stream = nullptr;
  hipEvent_t solved = nullptr;
  dv_batch_stats stats{};
  std::string err;
};

struct dv_server {
  struct Request {
    long ticket;
    dv_scene* scene;
    double bytes;
  };
  int build = DV_APGD_MIXED;
  int maxSlots = 0;
  double budget = 0, used = 0;
  dv::Engine engine;
  hipStream_t stream = nullptr;
  hipEvent_t solved = nullptr;
  int numCU = 0;
  std::deque<Request> waiting;
  std::vector<Request> slot;    // ticket < 0: free
  std::deque<long> completed;   // not yet reported
  std::deque<long> failed;      // diverged, not yet reported (dv_server_failed)
  long nextTicket = 1;
  std::string err;
};

namespace dv {
namespace {

template <class H, class F>
int guardedH(H* h, F&& fn) {
  if (!h) return DV_ERR_ARG;
  try {
    h->err.clear();
    return fn(*h);
  } catch (const HipError& e) {
    h->err = e.where + ": " + hipGetErrorString(e.code);
    return DV_ERR_HIP;
  } catch (const DvError& e) {
    h->err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    h->err = e.what();
    return DV_ERR_STATE;
  }
}

double msSince(const std::chrono::steady_clock::time_point& t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void needBuild(int build) {
  if (build != DV_APGD_FP64 && build != DV_APGD_MIXED)
    throw DvError{DV_ERR_UNSUPPORTED, "batches and servers run the graph-driven builds (FP64 or MIXED)"};
}

void needScene(dv_scene* h) {
  if (!h) throw DvError{DV_ERR_ARG, "null scene"};
  if (!h->s.initialized) throw DvError{DV_ERR_STATE, "scene not initialized"};
  if (h->s.antiRelax)
    throw DvError{DV_ERR_UNSUPPORTED, "anti-relaxation runs on the plain build, not in batches or servers"};
}

// Device memory the solver of a slot needs beyond the scene's own state:
// the FP64 iterate ring (4 slots + gammaHat), the product vectors (4 N gamma,
// r), J and JB, tmp, (mixed) the ring of D^T gamma, M^-1 and the FP32 frames,
// and the partial sums (per contact tile, per body tile and per heavy body).
double solverBytes(const Scene& s, bool mixed) {
  const double rows = 3.0 * std::max(s.ncCap, 1) + s.nbil;
  const double nc = std::max(s.ncCap, 1), nb = std::max(s.nb, 1);
  const double tm = mixed ? 4 : 8;
  return rows * (5 * 8 + 5 * tm) + nc * (8 * tm + (mixed ? 24 : 0)) + nb * ((mixed ? 21 : 4) * tm) +
         4.0 * 8 * 8 * (1 + (rows / 3 + 1023) / 1024) + 2 * 8 * (nb / 256 + 2 * nc / kHeavyDegree + 2) + 256;
}

// f(i) for i in [0, n) on up to eight host threads (each scene has its own
// stream and buffers, and the HIP runtime takes launches from several
// threads), so that the per-scene launch and read-back latencies of a
// batch overlap.  The first exception is rethrown after all threads end.
template <class F>
void parallelScenes(size_t n, F&& f) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const unsigned nt = (unsigned)std::min<size_t>(std::min(8u, hw), n / 16);
  if (nt <= 1) {
    for (size_t i = 0; i < n; ++i) f(i);
    return;
  }
  int dev = 0;
  DV_HIP(hipGetDevice(&dev));  // a new thread's current device is the runtime's default
  std::vector<std::thread> th;
  std::vector<std::exception_ptr> err(nt);
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      try {
        DV_HIP(hipSetDevice(dev));
        for (size_t i = t; i < n; i += nt) f(i);
      } catch (...) {
        err[t] = std::current_exception();
      }
    });
  for (std::thread& x : th) x.join();
  for (std::exception_ptr& e : err)
    if (e) std::rethrow_exception(e);
}

void batchStep(dv_batch& B) {
  if (B.scenes.empty()) return;
  const auto t0 = std::chrono::steady_clock::now();
  if (!B.stream) {
    DV_HIP(hipStreamCreateWithFlags(&B.stream, hipStreamNonBlocking));
    DV_HIP(hipEventCreateWithFlags(&B.solved, hipEventDisableTiming));
  }
  B.stats = dv_batch_stats{};
  std::vector<Scene*> solve;
  std::vector<char> rows(B.scenes.size());
  // all counts queued first
  parallelScenes(B.scenes.size(), [&](size_t i) { stepAssembleBegin(B.scenes[i]->s); });
  parallelScenes(B.scenes.size(), [&](size_t i) { rows[i] = stepAssembleEnd(B.scenes[i]->s); });
  for (size_t i = 0; i < B.scenes.size(); ++i) {
    Scene& s = B.scenes[i]->s;
    if (rows[i]) {
      DV_HIP(hipStreamWaitEvent(B.stream, s.ev[2], 0));
      solve.push_back(&s);
    }
  }
  const double msAsm = msSince(t0);
  const auto t1 = std::chrono::steady_clock::now();
  solveFusedMany(B.engine, solve, B.build == DV_APGD_MIXED, B.stream, B.scenes[0]->s.numCU,
                 B.scenes[0]->s.noGraph);
  DV_HIP(hipEventRecord(B.solved, B.stream));
  const double msSolve = msSince(t1);
  const auto t2 = std::chrono::steady_clock::now();
  for (size_t i = 0; i < B.scenes.size(); ++i) {  // a diverged solve fails the step before any update
    const Scene& s = B.scenes[i]->s;
    if (rows[i] && !std::isfinite(s.stats.residual))
      throw DvError{DV_ERR_DIVERGED, "APGD produced a non-finite residual"};
  }
  parallelScenes(B.scenes.size(), [&](size_t i) {  // all updates queued first
    Scene& s = B.scenes[i]->s;
    DV_HIP(hipStreamWaitEvent(s.stream, B.solved, 0));
    stepFinishQueue(s, rows[i] != 0, false);
  });
  parallelScenes(B.scenes.size(), [&](size_t i) { stepFinishWait(B.scenes[i]->s); });
  for (size_t i = 0; i < B.scenes.size(); ++i) {
    const Scene& s = B.scenes[i]->s;
    B.stats.total_contacts += s.nc;
    B.stats.total_iterations += s.stats.iterations;
    B.stats.max_iterations = std::max(B.stats.max_iterations, s.stats.iterations);
  }
  B.stats.num_scenes = (int)B.scenes.size();
  B.stats.rows_per_thread = B.engine.sched.rpt;
  B.stats.xcd_order = B.engine.sched.xcd;
  B.stats.wg_per_cu_contacts = B.engine.sched.wgc;
  B.stats.wg_per_cu_bodies = B.engine.sched.wgb;
  B.stats.schedule_tuning = B.engine.tuning;
  B.stats.ms_assembly = msAsm;
  B.stats.ms_solve = msSolve;
  B.stats.ms_update = msSince(t2);
  B.stats.ms_total = msSince(t0);
}

void serverOpen(dv_server& S, int build, int maxSlots, double budget) {
  needBuild(build);
  if (maxSlots < 2) throw DvError{DV_ERR_ARG, "a server needs at least two slots"};
  S.build = build;
  S.maxSlots = maxSlots;
  int dev = 0;
  DV_HIP(hipGetDevice(&dev));
  DV_HIP(hipDeviceGetAttribute(&S.numCU, hipDeviceAttributeMultiprocessorCount, dev));
  if (budget <= 0) {
    size_t freeB = 0, totalB = 0;
    DV_HIP(hipMemGetInfo(&freeB, &totalB));
    budget = 0.8 * (double)freeB;
  }
  S.budget = budget;
  DV_HIP(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
  DV_HIP(hipEventCreateWithFlags(&S.solved, hipEventDisableTiming));
  S.slot.assign(maxSlots, dv_server::Request{-1, nullptr, 0});
  engineOpen(S.engine, maxSlots, build == DV_APGD_MIXED, S.numCU, false, 1);
}

bool inFlight(const dv_server& S, const dv_scene* h) {
  for (const auto& r : S.waiting)
    if (r.scene == h) return true;
  for (const auto& r : S.slot)
    if (r.ticket >= 0 && r.scene == h) return true;
  return false;
}

// One turn of the server: launch a chunk of passes over the slots; while it
// runs, admit waiting requests (their collision detection and assembly run
// on their own streams, their setup is ordered after the chunk); then poll
// the slots and complete the requests whose solves are done.
void serverPoll(dv_server& S) {
  if (engineBusy(S.engine)) engineRunChunk(S.engine, S.stream);
  // admit in submission order while a slot is free and the budget allows
  // (an idle server always admits one request); the solver memory of a
  // request is estimated from its scene's capacities before its assembly,
  // so the requests are chosen first and then assembled together, from the
  // host threads
  std::vector<std::pair<dv_server::Request, int>> take;  // request, slot
  double used = S.used;
  const bool busy = engineBusy(S.engine);
  while (!S.waiting.empty()) {
    int free = -1;
    for (int k = 0; k < S.maxSlots && free < 0; ++k) {
      bool taken = S.slot[k].ticket >= 0;
      for (const auto& t : take) taken = taken || t.second == k;
      if (!taken) free = k;
    }
    if (free < 0) break;
    dv_server::Request r = S.waiting.front();
    r.bytes = solverBytes(r.scene->s, S.build == DV_APGD_MIXED);
    if (used + r.bytes > S.budget && (busy || !take.empty())) break;
    S.waiting.pop_front();
    used += r.bytes;
    take.emplace_back(r, free);
  }
  std::vector<char> rows(take.size());
  try {
    parallelScenes(take.size(), [&](size_t i) { stepAssembleBegin(take[i].first.scene->s); });
    parallelScenes(take.size(), [&](size_t i) { rows[i] = stepAssembleEnd(take[i].first.scene->s); });
  } catch (...) {
    // nothing was admitted: the taken requests wait again, in their order
    for (size_t i = take.size(); i-- > 0;) S.waiting.push_front(take[i].first);
    throw;
  }
  bool admitted = false;
  for (size_t i = 0; i < take.size(); ++i) {
    const dv_server::Request& r = take[i].first;
    Scene& s = r.scene->s;
    if (!rows[i]) {  // nothing to solve: the update completes the request
      stepFinish(s, false, false);
      S.completed.push_back(r.ticket);
      continue;
    }
    DV_HIP(hipStreamWaitEvent(S.stream, s.ev[2], 0));
    engineAdmit(S.engine, take[i].second, s);
    S.slot[take[i].second] = r;
    S.used += r.bytes;
    admitted = true;
  }
  if (admitted) engineStartAdmitted(S.engine, S.stream);
  // retire: gamma := gammaHat, stats, then each scene's update
  std::vector<int> done = enginePoll(S.engine, S.stream);
  if (done.empty()) return;
  engineRetire(S.engine, done, S.stream);
  DV_HIP(hipEventRecord(S.solved, S.stream));
  // every retired slot is freed; a solve with a non-finite residual fails
  // its request (no update: the scene keeps its state before the step)
  std::vector<dv_server::Request> ok;
  std::string failed;
  for (int k : done) {
    const dv_server::Request r = S.slot[k];
    S.slot[k] = dv_server::Request{-1, nullptr, 0};
    S.used -= r.bytes;
    const Scene& s = r.scene->s;
    if (!std::isfinite(s.stats.residual)) {
      S.failed.push_back(r.ticket);
      failed += (failed.empty() ? "" : ", ") + std::to_string(r.ticket);
    } else {
      ok.push_back(r);
    }
  }
  for (const auto& r : ok) {  // all updates queued first
    Scene& s = r.scene->s;
    DV_HIP(hipStreamWaitEvent(s.stream, S.solved, 0));
    stepFinishQueue(s, true, false);
  }
  for (const auto& r : ok) {
    stepFinishWait(r.scene->s);
    S.completed.push_back(r.ticket);
  }
  if (!failed.empty())
    throw DvError{DV_ERR_DIVERGED, "APGD produced a non-finite residual for request(s) " + failed};
}

}  // namespace
}  // namespace dv

extern "C" {

int dv_batch_create(dv_batch** out, int build) {
  if (!out) return DV_ERR_ARG;
  *out = nullptr;
  if (build != DV_APGD_FP64 && build != DV_APGD_MIXED) return DV_ERR_UNSUPPORTED;
  try {
    dv_batch* b = new dv_batch();
    b->build = build;
    *out = b;
    return DV_OK;
  } catch (...) {
    return DV_ERR_NOMEM;
  }
}

int dv_batch_destroy(dv_batch* b) {
  if (!b) return DV_ERR_ARG;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  b->engine.release();
  if (b->solved) (void)hipEventDestroy(b->solved);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return DV_OK;
}

const char* dv_batch_last_error(const dv_batch* b) { return b ? b->err.c_str() : "null batch"; }

int dv_batch_add(dv_batch* b, dv_scene* s) {
  return dv::guardedH(b, [&](dv_batch& B) {
    dv::needScene(s);
    for (dv_scene* x : B.scenes)
      if (x == s) throw dv::DvError{DV_ERR_ARG, "scene already in the batch"};
    B.scenes.push_back(s);
    return DV_OK;
  });
}

int dv_batch_step(dv_batch* b) {
  return dv::guardedH(b, [&](dv_batch& B) {
    dv::batchStep(B);
    return DV_OK;
  });
}

int dv_batch_get_stats(const dv_batch* b, dv_batch_stats* out) {
  if (!b || !out) return DV_ERR_ARG;
  *out = b->stats;
  return DV_OK;
}

int dv_server_create(dv_server** out, int build, int max_slots, double hbm_budget) {
  if (!out) return DV_ERR_ARG;
  *out = nullptr;
  dv_server* srv = new (std::nothrow) dv_server();
  if (!srv) return DV_ERR_NOMEM;
  const int rc = dv::guardedH(srv, [&](dv_server& S) {
    dv::serverOpen(S, build, max_slots, hbm_budget);
    return DV_OK;
  });
  if (rc != DV_OK) {
    g_createError = srv->err;
    dv_server_destroy(srv);
    return rc;
  }
  *out = srv;
  return DV_OK;
}

int dv_server_destroy(dv_server* srv) {
  if (!srv) return DV_ERR_ARG;
  if (srv->stream) (void)hipStreamSynchronize(srv->stream);
  srv->engine.release();
  if (srv->solved) (void)hipEventDestroy(srv->solved);
  if (srv->stream) (void)hipStreamDestroy(srv->stream);
  delete srv;
  return DV_OK;
}

const char* dv_server_last_error(const dv_server* srv) { return srv ? srv->err.c_str() : g_createError.c_str(); }

int dv_server_submit(dv_server* srv, dv_scene* scene, long* ticket) {
  return dv::guardedH(srv, [&](dv_server& S) {
    dv::needScene(scene);
    if (dv::inFlight(S, scene)) throw dv::DvError{DV_ERR_STATE, "scene already has a request in flight"};
    const long t = S.nextTicket++;
    S.waiting.push_back(dv_server::Request{t, scene, 0});
    if (ticket) *ticket = t;
    return DV_OK;
  });
}

int dv_server_poll(dv_server* srv, long* done, int max_done, int* n_done) {
  if (n_done) *n_done = 0;
  const int rc = dv::guardedH(srv, [&](dv_server& S) {
    dv::serverPoll(S);
    return DV_OK;
  });
  if (!srv) return rc;
  // completed requests are reported whether or not the poll failed
  int n = 0;
  while (!srv->completed.empty() && n < max_done && done) {
    done[n++] = srv->completed.front();
    srv->completed.pop_front();
  }
  if (n_done) *n_done = n;
  return rc;
}

int dv_server_failed(dv_server* srv, long* tickets, int max_n, int* n_out) {
  if (!srv || !n_out) return DV_ERR_ARG;
  int n = 0;
  while (!srv->failed.empty() && n < max_n && tickets) {
    tickets[n++] = srv->failed.front();
    srv->failed.pop_front();
  }
  *n_out = n;
  return DV_OK;
}

int dv_server_pending(const dv_server* srv, int* n) {
  if (!srv || !n) return DV_ERR_ARG;
  int k = (int)srv->waiting.size() + (int)srv->completed.size();
  for (const auto& r : srv->slot)
    if (r.ticket >= 0) ++k;
  *n = k;
  return DV_OK;
}

}  // extern "C"
