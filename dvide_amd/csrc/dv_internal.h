// Internal declarations of libdvide_hip.so: the device-resident scene and the
// host-side launchers of the collision, assembly, solver and update kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dvide.h"

namespace dv {

// Throwing HIP check used inside the library; the C ABI converts to codes.
struct HipError {
  hipError_t code;
  std::string where;
};
struct DvError {
  int code;
  std::string msg;
};
#define DV_HIP(call)                                                         \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) throw ::dv::HipError{e_, std::string(#call)};      \
  } while (0)

// A growable device array.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;
  void reserve(size_t n) {
    if (n <= cap) return;
    size_t want = n < 1024 ? 1024 : n + n / 2;
    T* q = nullptr;
    if (hipMalloc(&q, want * sizeof(T)) != hipSuccess)
      throw DvError{DV_ERR_NOMEM, "device allocation of " + std::to_string(want * sizeof(T)) + " bytes failed"};
    if (p) {
      (void)hipFree(p);
    }
    p = q;
    cap = want;
  }
  // grow keeping the first `keep` elements (stream-ordered copy)
  void reserveKeep(size_t n, size_t keep, hipStream_t st) {
    if (n <= cap) return;
    size_t want = n + n / 2;
    T* q = nullptr;
    if (hipMalloc(&q, want * sizeof(T)) != hipSuccess)
      throw DvError{DV_ERR_NOMEM, "device allocation failed"};
    if (p && keep) DV_HIP(hipMemcpyAsync(q, p, keep * sizeof(T), hipMemcpyDeviceToDevice, st));
    DV_HIP(hipStreamSynchronize(st));
    if (p) (void)hipFree(p);
    p = q;
    cap = want;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Bodies with more incidences than this are reduced by a whole workgroup.
constexpr int kHeavyDegree = 128;

constexpr int kSlots = 8;  // reduction slots of the solver kernels

// Most contact points a shell mesh may add (elements x (points per side - 2)^2).
constexpr long kMaxShellPoints = 1L << 26;
// Most bodies (shell contact points included) and contacts of one scene: row
// and DOF offsets (3 per body or contact, 2 incidences per contact) are ints.
constexpr long kMaxBodies = 1L << 28;
constexpr long kMaxContacts = 1L << 28;

struct ApgdState;
struct SolveDims;
struct Scene;

// An instantiated hipGraph and the launch parameters it was captured with.
// The stream it was last launched on is kept, so that an exec is destroyed
// only after its last launch has finished.
struct GraphCache {
  hipGraphExec_t exec = nullptr;
  hipStream_t launchedOn = nullptr;
  std::vector<char> key;
  void reset();
  bool matches(const void* k, size_t n) const;
  void setKey(const void* k, size_t n);
  // instantiates `graph` (destroyed here) as the cached exec; the cache is
  // left empty if instantiation fails
  void instantiate(hipGraph_t graph, const char* what);
  void launch(hipStream_t st);
};

// A tile of solver work: rows (or bodies) [start, start + tile size) of one
// scene of an engine's table.
struct Tile {
  int scene, start;
};

// Tiles of a set of scene slots: row tiles, body tiles, the heavy bodies and
// their counts (device copies), and the slot list.
struct TileSet {
  DevBuf<Tile> items, bodies;
  DevBuf<int2> heavy;
  DevBuf<int> heavyIdx;  // each heavy body's index in its slot's heavy list
  DevBuf<int> counts;  // [0] row tiles, [1] body tiles, [2] heavy bodies
  DevBuf<int> slots;
  std::vector<Tile> hItems, hBodies;
  std::vector<int> hSlots;
  int countsHost[4] = {};
  void release();
};

// Launch schedule of the solver's pass kernels.  None of its choices changes
// a result: rows per thread only moves rows between lanes (the partial sums
// keep the one-row-per-thread order), and the tile order and grid sizes only
// decide which workgroup takes which tile.
struct Schedule {
  int rpt = 1;  // contact rows per thread of P1 (1 or 2)
  int xcd = 1;  // XCD-contiguous tile order for the gathering kernels
  int wgc = 32; // several slots: contact-kernel workgroups per CU
  int wgb = 32; // several slots: body-kernel workgroups per CU (one slot: 4)
};

// The graph-driven APGD solver over a table of scene slots
// (dv_apgd_fused.hip): the device-resident table, the tile sets of the
// active, newly admitted and retiring slots, the per-slot states the host
// polls, and the captured graph of a chunk of passes.  Slots join (after
// their setup) and leave (after their finalisation) between chunks.
struct Engine {
  int mixed = -1;       // 1: FP32 products; 0: FP64
  Schedule sched;       // of the chunk being launched
  int forcedRpt = 0;    // 1 or 2: a fixed schedule with these rows per thread; 0: tuned
  int tuning = 0;       // the last solve measured candidate schedules
  hipEvent_t tev[2][2] = {};  // timing of each in-flight chunk while tuning
  int chunkCand[2] = {-1, -1};  // candidate schedule of each in-flight chunk (-1: none)
  std::string tuneKey;  // the class being measured
  int nSlots = 0;
  std::vector<char> hostTable;  // Args<TM>[nSlots]
  std::vector<Scene*> slotScene;
  DevBuf<char> table;
  TileSet active, admit, retire;
  DevBuf<char> out;     // ApgdState[nSlots]
  DevBuf<char> states;  // ApgdState[nSlots][2], by pass parity
  DevBuf<char> views;   // PassView[nSlots][2]: what the pass kernels read of the states
  void* outHost = nullptr;  // pinned copies of out: two halves, by chunk parity
  hipEvent_t polled[2] = {nullptr, nullptr};  // after each half's copy
  int outHostCap = 0;
  GraphCache graph;
  GraphCache burst;     // the timed burst of one pass kernel (DV_PARAM_PROFILE)
  int numCU = 0;
  bool noGraph = false;
  long launched = 0;    // passes launched since the last reset
  void release();
};

struct Scene {
  // --- parameters (System.cu:14-28, APGD.cu:6-16, CollisionDetector.cu:503-514)
  double h = 1e-3, mu = 0.25, tol = 1e-4, envelope = 1e-3;
  int maxIt = 100000;
  double g[3] = {0, -9.81, 0};
  int bins[3] = {20, 20, 20};
  bool warm = false;
  bool antiRelax = false;   // APGD::useAntiRelaxation (APGD.cu:15)
  int build = DV_APGD_FP64; // DV_PARAM_APGD_BUILD
  int geomMode = 1;
  int verbose = 0;
  int grid = 0;
  bool noGraph = false;  // launch the solver passes directly (profiling)
  int profKernel = 0;    // DV_PARAM_PROFILE
  int rowsPerThread = 0; // DV_PARAM_ROWS_PER_THREAD (0: the tuned choice)
  std::vector<hipEvent_t> profEv;
  double time = 0;
  int timeIndex = 0;
  bool initialized = false;
  std::string err;

  // --- host staging (System::bodies and the constraint lists)
  std::vector<double> hPos, hVel, hMass, hGeom;
  std::vector<int> hFixed, hFamily;
  std::vector<int> hBilA, hBilB;
  std::vector<double> hBilVel, hBilT0, hBilP0;

  // --- fixed ANCF shell mesh (System::importMesh, System.cu:1932-2060).  Its
  // contact points (System.cu:494-499) join the bodies at initialisation as
  // fixed spheres of half the thickness in collision family -2, after the
  // caller's bodies and in the reference's collision-geometry order, so
  // collision detection, assembly and the solvers treat them as bodies; the
  // boundary maps them back onto the shell's 9 DOFs per node.
  int nUser = 0;                    // the caller's bodies (from initialisation)
  int nNodes = 0;
  std::vector<double> hShellNodes;  // 9 per node: position, x-gradient, y-gradient
  std::vector<double> hShellFext;   // 9 per node
  std::vector<int> hShellConn;      // 4 node indices per element
  std::vector<double> hShellGeom;   // 3 per element: a, b, thickness
  int shellPoints = 0;              // contact points per element side
  std::vector<int> hPtEl;           // per contact point: its element
  std::vector<double> hPtS;         // per contact point: 12 shape-function weights
  DevBuf<double> shellNodes, ptS, shellF;
  DevBuf<int> shellConn, ptEl, nodeIncStart, nodeInc;

  // --- device state
  hipStream_t stream = nullptr;
  int device = 0, numCU = 0;
  int nb = 0, nbil = 0;
  int nsmall = 0, nbig = 0;
  double cellSize = 0;
  unsigned hashSize = 0;
  DevBuf<double> p, v, f, k, fContact, tmp, minv;  // minv per body
  DevBuf<double> geom;                             // 3 per body
  DevBuf<int> family, smallIds, bigIds;
  // bilateral rows
  DevBuf<int> bilA, bilB;
  DevBuf<double> bilVel, bilT0, bilP0;
  DevBuf<int> bilIncStart, bilInc;  // per body: (row<<3)|(comp<<1)|sign

  // --- collision detection work
  DevBuf<unsigned> slotOf, slotCount, slotStart, slotFill;
  DevBuf<int> cellItems;
  DevBuf<double> cellPosR;  // 4 per sorted small geometry
  DevBuf<int> cntUp, cntDown, cStart, iStart, scanTmp;
  DevBuf<int> counters;     // [0] nc, [1] ninc, [2] nHeavy
  DevBuf<int> hostTotals;
  int* pinned = nullptr;    // pinned host copy of counters

  // --- contacts (canonical (idA, idB) order)
  int nc = 0, ncCap = 0;
  DevBuf<int> cA, cB;
  DevBuf<double> frame;  // 9 SoA arrays of ncCap: n, u, v components
  DevBuf<double> gap;    // w = -penetration
  DevBuf<double> loc;    // 3 SoA arrays of ncCap
  DevBuf<float> frame32; // FP32 copy for the mixed solver
  DevBuf<int> inc;       // per-body incidences (contact << 1 | side)
  DevBuf<int> posB;      // per contact: its slot in the B-side order (see buildIncidence)
  DevBuf<int> heavy;     // bodies processed by a whole workgroup, ascending
  DevBuf<int> heavyPos;  // their positions in `heavy` (exclusive scan of the flags)
  DevBuf<int> bigChunks; // per (big geometry, chunk) partner counts

  // --- system of the step (flat rows: nbil bilateral, then 3 per contact)
  int R = 0, gammaLen = 0;
  DevBuf<double> r, b, gamma;
  // fused solver work
  DevBuf<double> work;   // solver iterates (FP64 in every solver)
  DevBuf<double> workM64;  // products of the FP64 solver
  DevBuf<float> work32;  // products and matrix data of the mixed solver
  DevBuf<double> partials;
  Engine engine;               // the graph-driven solver of this scene alone

  // --- statistics of the last step
  dv_stats stats{};
  hipEvent_t ev[6] = {};
};

// collision detection + incidence lists (dv_collide.hip)
void collideInit(Scene& s);
void collide(Scene& s);
// the two halves of collide: counts (queued, no wait), then the wait for
// them and the emission of the contacts; many scenes can queue their counts
// before the first wait
void collideCount(Scene& s);
void collideEmit(Scene& s);

// fixed ANCF shells (dv_shell.hip): the contact points appended to the
// host body arrays, their positions computed on the device, and the contact
// force on the shell DOFs (9 per node, into a device buffer of 9 nNodes)
void shellAppendPoints(Scene& s);
void shellPlacePoints(Scene& s);
void shellForces(Scene& s, double* out);

// assembly of k and r, update (dv_assemble.hip)
void buildK(Scene& s);
void buildR(Scene& s);
void updateState(Scene& s, bool haveRows, bool exactOrder);
void schurExact(Scene& s, const double* x, double* y);  // device pointers, flat R
void gatherExact(Scene& s, const double* x, double* fOut, double* tmpOut, bool skipFixed);
void ensureRowBuffers(Scene& s);
void fillDevice(Scene& s, double* x, long n, double v);
void rowsD(Scene& s, const double* x, const double* badd, double* y);
void minvTimes(Scene& s, const double* x, double* y);  // y = M^-1 x per DOF

// solvers (dv_apgd_plain.hip, dv_apgd_fused.hip)
void solvePlain(Scene& s);
void solveFused(Scene& s, bool mixed);
// N x by the graph-driven solver's own pass kernels (J store, P2, contact rows) on the
// assembled scene; x is a flat device array, y a flat host array
void passProduct(Scene& s, bool mixed, const double* xDev, double* yHost);
// one solve of several scenes in the same launches (on stream st); each
// scene's results equal those of solveFused on it alone
void solveFusedMany(Engine& E, const std::vector<Scene*>& scenes, bool mixed, hipStream_t st, int numCU,
                    bool noGraph);
// continuous batching on an engine (dv_apgd_fused.hip): slots are admitted
// (scene assembled, setup launched) and retired (gamma written back, stats
// filled) between chunks of passes
// forcedRpt: 1 or 2 fixes the schedule (with that many rows per thread);
// 0 lets the schedule tuner choose
void engineOpen(Engine& E, int nSlots, bool mixed, int numCU, bool noGraph, int forcedRpt);
void engineAdmit(Engine& E, int slot, Scene& s);
void engineStartAdmitted(Engine& E, hipStream_t st);
void engineRunChunk(Engine& E, hipStream_t st);
// copies the slots' states to the host; returns the active slots now done
std::vector<int> enginePoll(Engine& E, hipStream_t st);
void engineRetire(Engine& E, const std::vector<int>& slots, hipStream_t st);
bool engineBusy(const Engine& E);

// scans (dv_scan.hip): exclusive scan of n ints, out has n+1 entries
void exclusiveScan(Scene& s, const int* in, int* out, int n);
void exclusiveScanU(Scene& s, const unsigned* in, unsigned* out, int n);

inline int blocksFor(long n, int t) { return (int)((n + t - 1) / t) > 0 ? (int)((n + t - 1) / t) : 1; }

}  // namespace dv
