// APGD (APGD::solve, APGD.cu:164-361) as a device-resident state machine
// driven by a captured hipGraph: no host round trip per reduction.  One
// engine solves one scene or a batch of independent scenes in the same
// launches.
//
// One solver pass (one Schur product plus the vector algebra around it) is
// two kernels over all active scenes:
//   P1 (contacts): N gamma_k = F (tmp_A - tmp_B) from the previous pass's
//       tmp when the previous decision accepted gamma_k (else read back),
//       y and N y from the last two iterates, g = N y + r,
//       gamma_new = Pi(y - t g), the dot products of the objective and
//       restart tests, J_c = F_c^T gamma_new (the world-frame impulse of
//       contact c on body A), and the infinity norm of getResidual
//       (APGD.cu:90-102) of the iterate the previous decision accepted;
//   P2 (bodies):   s_b = sum_A J_c - sum_B J_c + bilateral rows (D^T
//       gamma_new), tmp_b = M_b^-1 s_b, and the Lipschitz test's body terms:
//       s . tmp, whose sum is gamma_new . N gamma_new (FP64), or, with the
//       ring of per-body D^T gamma, d . M^-1 d for d = D^T (gamma_new - y),
//       whose sum is Delta' N Delta (mixed); bodies with many contacts are
//       reduced by a whole workgroup;
// followed by each scene's decision: a fixed-order reduction of its partial
// sums, the Lipschitz test (APGD.cu:236-264) and either a backtracking step
// or the acceptance logic (APGD.cu:266-332).  Several scenes decide in a
// kernel of their own (one workgroup per scene); a single scene's decision
// is taken by every workgroup of the next pass's P1, from the same totals in
// the same order, and the decide kernel runs once per chunk of passes.
//
// Work is cut into tiles of one scene (kTC rows or kTB bodies); a scene's
// partial sums are kept per tile and reduced in tile order, so a scene gives
// bit-identical results alone and in any batch.  State and partials are
// double-buffered by pass parity, so no decision races with its
// publication.  Kernel boundaries are the grid-wide barriers: on MI355X a
// kernel boundary inside a graph costs ~1.6 us, less than a grid barrier of
// a persistent cooperative kernel (3.7 us at 128 workgroups, ~12 us at 512).
// The host launches the graph of a chunk of passes, reads the scenes'
// states, and repeats; scenes that have converged skip their tiles.
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
// difference obj2 - obj1; D'N D = |M^-1/2 D^T D|^2 is summed over the bodies.
#include "dv_geom.h"
#include "dv_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <sys/stat.h>

namespace dv {

struct ApgdState {
  double L, t, theta, beta, best, objective;
  int it, bt, backtracks, restart;
  int cur, prev, nxt;
  int hatSrc, done, converged, passes;
  int pendingRes;  // the iterate in slot cur was accepted and its residual is still to be taken
  int fresh;       // N of ring slot cur is still to be formed (by the next P1, from tmp)
};

// What the pass kernels read of a slot's state, written beside the state by
// every decision.  The ring slots, the done and restart flags come first, in
// one 16-byte load; the ring pointers follow by arithmetic from the slot's
// ring base, so a pass kernel's data loads wait on a single round trip.
struct alignas(32) PassView {
  int cur, prev, nxt, flags;  // flags: 1 done, 2 restart, 4 fresh
  double t, beta;
};
constexpr int kViewDone = 1, kViewRestart = 2, kViewFresh = 4;

__device__ __forceinline__ PassView viewOf(const ApgdState& S) {
  PassView v;
  v.cur = S.cur; v.prev = S.prev; v.nxt = S.nxt;
  v.flags = (S.done ? kViewDone : 0) | (S.restart ? kViewRestart : 0) | (S.fresh ? kViewFresh : 0);
  v.t = S.t; v.beta = S.beta;
  return v;
}

struct SolveDims {
  int nc, nbil, nb, maxIt;
  double mu, tol, gdiff, igdiff;  // igdiff = 1 / gdiff (APGD.cu:98-100)
  int warm, geomMode;
};

namespace {

constexpr int kTC = 512;   // contact kernels: 8 wavefronts, one tile of rows
constexpr int kTB = 256;   // body kernel: one tile of bodies
constexpr int kTD = 64;    // setup and final decisions: one wavefront per scene
constexpr int kBacktrackCap = 1100;  // L doubles past 1e308 long before this
constexpr int kPassesPerChunk = 16;  // even: a chunk ends on parity 0
static_assert(kPassesPerChunk % 2 == 0, "pass parity must return to 0 at a chunk end");

template <typename TM>
struct alignas(4 * sizeof(TM)) V4 {
  TM x, y, z, w;
};

// Device code reaches the solver's arrays through pointers held in the scene
// table, which the compiler can only treat as generic (flat) pointers.  Flat
// loads also count against the LDS/scalar counter, so every wait for a
// scalar load of the table would also wait for all outstanding data loads.
// Casting to the global address space makes them global loads and stores.
#define DV_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const DV_GLOBAL T* gl(const T* p) {
  return (const DV_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ DV_GLOBAL T* gl(T* p) {
  return (DV_GLOBAL T*)p;
}
template <typename TM>
__device__ __forceinline__ V4<TM> ld4(const V4<TM>* p, long i) {
  const DV_GLOBAL V4<TM>* q = gl(p) + i;
  V4<TM> v;
  v.x = q->x; v.y = q->y; v.z = q->z; v.w = q->w;
  return v;
}
template <typename TM>
__device__ __forceinline__ void st4(V4<TM>* p, long i, const V4<TM>& v) {
  DV_GLOBAL V4<TM>* q = gl(p) + i;
  q->x = v.x; q->y = v.y; q->z = v.z; q->w = v.w;
}

// Per-scene view of the solver, one entry of the device-resident scene table.
template <typename TM>
struct Args {
  SolveDims dims;
  int nTiles;           // row tiles of this solve (partial sums per slot)
  int active;           // the slot holds a scene being solved
  const double* frame64;  // assembled frames: 9 SoA arrays n, u, v of stride cap
  TM* frame;            // frames the passes read: frame64's n (FP64) or an FP32 n, u copy
  long cap;             // contact capacity: row-vector stride
  const int* cA;
  const int* cB;
  const int* iStart;
  const int* cStart;    // contacts of body b as body A: [cStart[b], cStart[b+1]) (rows sorted by idA)
  const int* posB;      // slot of contact c in the B-side order, JB[posB[c]] = J[c]
  const int* bilIncStart;
  const int* bilInc;
  const int* bilA;
  const int* bilB;
  const int* heavy;
  const int* counters;  // [2] = number of heavy bodies
  const double* minv64;
  TM* minv;             // minv64 (FP64) or its FP32 copy
  const double* rFlat;  // assembled r, reference layout
  TM* r;                // r, component-major
  double* gammaIO;      // flat gamma: previous values in, result out
  double* ring[4];      // iterates (component-major), slots named by ApgdState
  TM* rn[4];            // N of the iterate in the same ring slot
  long ringStride;      // ring[k] = ring[0] + k * ringStride, and the same for rn
  V4<TM>* J;            // per contact, in contact order: body A's incidences are contiguous
  V4<TM>* JB;           // the same, in B-side order: body b's B incidences are
                        // [iStart[b] - cStart[b], iStart[b+1] - cStart[b+1])
  V4<TM>* tmp;
  V4<TM>* sr;           // mixed solver: D^T of each ring slot's iterate, per body (slot k at k * nb)
  double* gHat;         // gammaHat before any iterate is accepted (ones)
  double* part1[2];     // kSlots per tile, by pass parity
  double* part3[2];     // setup and final kernels: a second kSlots per tile
  double* lq[2];        // P2's Lipschitz partials, by pass parity: one per body tile, then one
                        // per heavy body (in the heavy list's order)
  double* p1sum[2];     // single slot: P1's slots totalled by P2's last workgroup
  int nBodyTiles;       // body tiles of the scene (kTB bodies each)
};

// Launch-wide tables (device memory, so that a captured graph stays valid
// while their contents change).
struct Tables {
  const Tile* items;    // row tiles (kTC rows of one slot)
  const Tile* bodies;   // body tiles (kTB bodies of one slot)
  const int2* heavy;    // (slot, body) of bodies reduced by a workgroup
  const int* heavyIdx;  // the index of each in its slot's heavy list
  const int* counts;    // [0] row tiles, [1] body tiles, [2] heavy bodies
  const int* slots;     // the set's slots (one decide workgroup each)
  ApgdState* st;        // [slot][2]: the solver state by pass parity
  PassView* pv;         // [slot][2]: its view for the pass kernels
  ApgdState* out;       // per slot: the state after the latest decision
  int single;           // one slot: tile k is rows k*kTC.. of slot 0 (no table read)
  int xcd;              // XCD-contiguous tile order for the gathering kernels
};

// state of slot `slot` at parity k: at a fixed offset from the kernel
// argument, so reading it does not wait for a table load
#define STATE(T, slot, k) ((T).st[2 * (slot) + (k)])
#define VIEW(T, slot, k) ((T).pv[2 * (slot) + (k)])

// A slot's pass view.  viewDone tests the ring slots (never negative) with
// the done flag, so that the compiler reads the view's first 16 bytes in one
// load before the branch instead of the flag alone and the slots after it.
__device__ __forceinline__ PassView loadView(const Tables& T, int slot, int k) {
  return VIEW(T, slot, k);
}
__device__ __forceinline__ bool viewDone(const PassView& S) {
  return (S.flags & kViewDone) || (S.cur | S.prev | S.nxt) < 0;
}

// row offsets: contact c component k -> k*cap + c; bilateral b -> 3*cap + b
#define ROW(a, k, c) ((long)(k) * (a).cap + (c))
#define BROW(a, b) (3L * (a).cap + (b))

// ---------------------------------------------------------------------------
// reductions (slot kSlots-1 is a max, the others sums)

// Wavefront reductions without the LDS crossbar: within each row of 16
// lanes the partner values come by DPP moves (lanes l ^ 1, l ^ 2, then the
// mirrored lane in the half-row and in the row), so every lane holds its
// row's total; the four row totals are read from lanes 0, 16, 32, 48 and
// combined as (r0 + r1) + (r2 + r3).  The order is fixed, so the result is
// the same in every lane and on every run.  Callers run them with all 64
// lanes active (a disabled source lane would read as 0).  Against the LDS
// butterfly (six dependent ds_bpermute round trips) this took P2, whose extra
// workgroup reduces eight slots, from 6.6 us to 5.4 us.
template <int kCtrl>
__device__ __forceinline__ double dppMove(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, kCtrl, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), kCtrl, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double laneValue(double v, int lane) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
constexpr int kDppXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141; // row_half_mirror
constexpr int kDppMirror = 0x140;     // row_mirror
__device__ __forceinline__ double waveSum(double v) {
  v += dppMove<kDppXor1>(v);
  v += dppMove<kDppXor2>(v);
  v += dppMove<kDppHalfMirror>(v);
  v += dppMove<kDppMirror>(v);
  return (laneValue(v, 0) + laneValue(v, 16)) + (laneValue(v, 32) + laneValue(v, 48));
}
__device__ __forceinline__ double waveMax(double v) {
  v = fmax(v, dppMove<kDppXor1>(v));
  v = fmax(v, dppMove<kDppXor2>(v));
  v = fmax(v, dppMove<kDppHalfMirror>(v));
  v = fmax(v, dppMove<kDppMirror>(v));
  return fmax(fmax(laneValue(v, 0), laneValue(v, 16)), fmax(laneValue(v, 32), laneValue(v, 48)));
}

// Workgroup totals of NS slots over a tile whose threads hold R rows each:
// row u * NT + t of the tile is thread t's accumulator acc[u].  Each
// accumulator is reduced over its wavefront, and the wavefront totals are
// combined in row order, so the result is that of one row per thread on
// R * NT threads, bit for bit: the rows-per-thread schedule changes the
// work per lane, not the arithmetic.
//
// Slots outside kUsed are left out and given as 0: a caller whose build
// never adds to a slot saves its wavefront reduction (the mixed contact
// pass fills 3 of its 8 slots).
template <int NT, int R, int NS, unsigned kUsed = 0xFFFFFFFFu>
__device__ void blockReduceR(double (&acc)[R][NS], double* out) {
  constexpr int W = NT / 64;
  __shared__ double sh[R * W][NS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int u = 0; u < R; ++u) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      if (!((kUsed >> s) & 1u)) continue;
      double v = (s % kSlots == kSlots - 1) ? waveMax(acc[u][s]) : waveSum(acc[u][s]);
      if (lane == 0) sh[u * W + wave][s] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < NS) {
    const int s = threadIdx.x;
    double v = 0.0;
    if ((kUsed >> s) & 1u) {
      v = sh[0][s];
      for (int w = 1; w < R * W; ++w) v = (s % kSlots == kSlots - 1) ? fmax(v, sh[w][s]) : v + sh[w][s];
    }
    out[s] = v;
  }
  __syncthreads();
}

template <int NT, int NS>
__device__ void blockReduce(double (&acc)[NS], double* out) {
  blockReduceR<NT, 1, NS>(reinterpret_cast<double(&)[1][NS]>(acc), out);
}

template <int NT, int R = 1, unsigned kUsed = 0xFFFFFFFFu>
__device__ void writePartialsR(double (&acc)[R][kSlots], double* part) {
  __shared__ double tot[kSlots];
  blockReduceR<NT, R, kSlots, kUsed>(acc, tot);
  if (threadIdx.x < kSlots) gl(part)[threadIdx.x] = tot[threadIdx.x];
}

template <int NT>
__device__ void writePartials(double (&acc)[kSlots], double* part) {
  writePartialsR<NT, 1>(reinterpret_cast<double(&)[1][kSlots]>(acc), part);
}

// Fixed-order totals of the per-tile partials of one scene by one wavefront
// (all 64 lanes): out[0..7] from part1, out[8..15] from part3, valid in every
// lane.  Lane l adds tiles l, l+64, l+128, ... in order; the lanes are then
// combined by a butterfly.  Loads are issued four tiles at a time.
__device__ void reduceAll(const double* part1, const double* part3, int ng, double* out) {
  const int lane = threadIdx.x & 63;
  double v[2 * kSlots];
#pragma unroll
  for (int s = 0; s < 2 * kSlots; ++s) v[s] = 0.0;
  for (int base = 0; base < ng; base += 256) {
    double x[4][2 * kSlots];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int w = base + lane + 64 * k;
      const int ws = w < ng ? w : 0;
#pragma unroll
      for (int s = 0; s < kSlots; ++s) {
        x[k][s] = gl(part1)[(long)ws * kSlots + s];
        x[k][kSlots + s] = gl(part3)[(long)ws * kSlots + s];
      }
    }
#pragma unroll
    for (int s = 0; s < 2 * kSlots; ++s) {
      const bool isMax = (s % kSlots == kSlots - 1);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double xv = (base + lane + 64 * k) < ng ? x[k][s] : 0.0;
        v[s] = isMax ? fmax(v[s], xv) : v[s] + xv;
      }
    }
  }
#pragma unroll
  for (int s = 0; s < 2 * kSlots; ++s) out[s] = (s % kSlots == kSlots - 1) ? waveMax(v[s]) : waveSum(v[s]);
}

// Fixed-order totals over a slot's per-tile partials of NS consecutive slots
// from s0 (the last slot, kSlots - 1, is a max, the others sums), by a
// workgroup of kTP threads: thread t adds tiles t, t + kTP, ... in order (the
// loads of K tiles issued together; K sets only how many are in flight, not
// the order), each wavefront combines its lanes by
// a butterfly, and the wavefront totals are added in wavefront order.  out[]
// holds the NS totals in every thread.  The pass decision reads P1's slots
// and P2's Lipschitz partials this way, wherever it takes them from.  In a wider
// workgroup the threads past kTP add nothing, so the totals are the same.
constexpr int kTP = 256;
// Slots outside kUsed are not read and are given as 0.
template <int NS, int kStride = kSlots, int K = 8, unsigned kUsed = 0xFFFFFFFFu>
__device__ void reduceSlots(const double* part, int s0, int ng, double (&out)[NS]) {
  constexpr int W = kTP / 64;
  __shared__ double sh[W][NS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool mine = threadIdx.x < kTP;
  double v[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) v[s] = 0.0;
  for (int base = 0; base < ng; base += kTP * K) {
    double x[K][NS];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int w = base + (int)threadIdx.x + kTP * k;
      const bool on = mine && w < ng;
#pragma unroll
      for (int s = 0; s < NS; ++s)
        x[k][s] = on && ((kUsed >> s) & 1u) ? gl(part)[(long)w * kStride + s0 + s] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      if (!((kUsed >> s) & 1u)) continue;
      const bool isMax = s0 + s == kSlots - 1;
#pragma unroll
      for (int k = 0; k < K; ++k) v[s] = isMax ? fmax(v[s], x[k][s]) : v[s] + x[k][s];
    }
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    if (!((kUsed >> s) & 1u)) continue;
    const double t = s0 + s == kSlots - 1 ? waveMax(v[s]) : waveSum(v[s]);
    if (lane == 0 && wave < W) sh[wave][s] = t;
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    double t = 0.0;
    if ((kUsed >> s) & 1u) {
      t = sh[0][s];
      for (int w = 1; w < W; ++w) t = s0 + s == kSlots - 1 ? fmax(t, sh[w][s]) : t + sh[w][s];
    }
    out[s] = t;
  }
  __syncthreads();
}

// The slots of P1's partial sums each build adds to (the others stay zero).
template <typename TM>
constexpr unsigned kP1Slots = sizeof(TM) < sizeof(double) ? 0xB0u : 0xBFu;

// The totals the pass decision needs: P1's slots (kP1Slots: 4, 5 and 7 in
// the mixed build, 0..5 and 7 in the FP64 build) and, in slot kSlots + 6,
// P2's Lipschitz term (its partials per body tile and per heavy body, packed
// in lq, nl of them); the other slots are 0.  reducePass reduces P1's partials; passTotals takes them already
// totalled (a single slot's P2 totals them), which keeps the registers of
// the eight-slot reduction out of the kernels that decide a single slot.
template <typename TM>
__device__ void reducePass(const double* part1, int ng, const double* lq, int nl, double* out) {
#pragma unroll
  for (int s = 0; s < 2 * kSlots; ++s) out[s] = 0.0;
  double t1[kSlots], t3[1];
  reduceSlots<kSlots, kSlots, 8, kP1Slots<TM>>(part1, 0, ng, t1);
  reduceSlots<1, 1>(lq, 0, nl, t3);
#pragma unroll
  for (int s = 0; s < kSlots; ++s) out[s] = t1[s];
  out[kSlots + 6] = t3[0];
}
__device__ void passTotals(const double* p1, const double* lq, int nl, double* out) {
#pragma unroll
  for (int s = 0; s < 2 * kSlots; ++s) out[s] = 0.0;
  double t3[1];
#pragma unroll
  for (int s = 0; s < kSlots; ++s) out[s] = gl(p1)[s];
  reduceSlots<1, 1>(lq, 0, nl, t3);
  out[kSlots + 6] = t3[0];
}

// ---------------------------------------------------------------------------
// contact frames and row helpers

__device__ __forceinline__ void ldFrame(const Args<double>& a, int c, int mode, double (&F)[9]) {
  const long S = a.cap;
  const d3 n = mk3(gl(a.frame)[c], gl(a.frame)[S + c], gl(a.frame)[2 * S + c]);
  d3 u, v;
  contactFrame(n, mode, &u, &v);  // bit-identical to the assembled frame
  F[0] = n.x; F[1] = n.y; F[2] = n.z;
  F[3] = u.x; F[4] = u.y; F[5] = u.z;
  F[6] = v.x; F[7] = v.y; F[8] = v.z;
}

__device__ __forceinline__ void ldFrame(const Args<float>& a, int c, int, float (&F)[9]) {
  const long S = a.cap;
#pragma unroll
  for (int k = 0; k < 6; ++k) F[k] = gl(a.frame)[k * S + c];
  F[6] = F[1] * F[5] - F[2] * F[4];  // v = n x u
  F[7] = F[2] * F[3] - F[0] * F[5];
  F[8] = F[0] * F[4] - F[1] * F[3];
}

// J_c = F_c^T x_c, stored twice: at c (contact order, where body A's
// incidences are contiguous) and at posB[c] (B-side order, where body B's
// are), so that P2 reads each body's incidences as two contiguous runs
template <typename TM>
__device__ __forceinline__ void storeJ(const Args<TM>& a, int c, int cb, const TM (&F)[9], TM x0, TM x1, TM x2) {
  V4<TM> j;
  j.x = F[0] * x0 + F[3] * x1 + F[6] * x2;
  j.y = F[1] * x0 + F[4] * x1 + F[7] * x2;
  j.z = F[2] * x0 + F[5] * x1 + F[8] * x2;
  j.w = 0;
  st4(a.J, c, j);
  st4(a.JB, cb, j);
}
template <typename TM>
__device__ __forceinline__ void storeJ(const Args<TM>& a, int c, const TM (&F)[9], TM x0, TM x1, TM x2) {
  storeJ(a, c, gl(a.posB)[c], F, x0, x1, x2);
}

template <typename TM>
__device__ __forceinline__ TM tmpDof(const Args<TM>& a, int dof) {
  const V4<TM> t = ld4(a.tmp, dof / 3);
  const int k = dof % 3;
  return k == 0 ? t.x : (k == 1 ? t.y : t.z);
}

template <typename TM>
__device__ __forceinline__ TM bilRow(const Args<TM>& a, int q) {
  TM s = tmpDof(a, gl(a.bilA)[q]);
  if (gl(a.bilB)[q] >= 0) s -= tmpDof(a, gl(a.bilB)[q]);
  return s;
}

// (N x)_c = F_c (tmp_A - tmp_B): the three rows of a contact from the two
// bodies' M^-1 D^T x
template <typename TM>
__device__ __forceinline__ void rowsOfTmp(const TM (&F)[9], const V4<TM>& tA, const V4<TM>& tB, TM& n0, TM& n1,
                                          TM& n2) {
  const TM d0 = tA.x - tB.x, d1 = tA.y - tB.y, d2 = tA.z - tB.z;
  n0 = F[0] * d0 + F[1] * d1 + F[2] * d2;
  n1 = F[3] * d0 + F[4] * d1 + F[5] * d2;
  n2 = F[6] * d0 + F[7] * d1 + F[8] * d2;
}

template <typename TM>
__device__ __forceinline__ void contactRows(const Args<TM>& a, int c, const TM (&F)[9], TM& n0, TM& n1, TM& n2) {
  const V4<TM> tA = ld4(a.tmp, gl(a.cA)[c]), tB = ld4(a.tmp, gl(a.cB)[c]);
  const TM d0 = tA.x - tB.x, d1 = tA.y - tB.y, d2 = tA.z - tB.z;
  n0 = F[0] * d0 + F[1] * d1 + F[2] * d2;
  n1 = F[3] * d0 + F[4] * d1 + F[5] * d2;
  n2 = F[6] * d0 + F[7] * d1 + F[8] * d2;
}

__device__ __forceinline__ double residualContact(double g0, double g1, double g2, double w0, double w1,
                                                  double w2, double mu, double gdiff, double ig) {
  double z0 = 1.0 * g0 + (-gdiff) * w0, z1 = 1.0 * g1 + (-gdiff) * w1, z2 = 1.0 * g2 + (-gdiff) * w2;
  projectCone<double>(z0, z1, z2, mu);
  const double e0 = ig * g0 + (-ig) * z0, e1 = ig * g1 + (-ig) * z1, e2 = ig * g2 + (-ig) * z2;
  // fmax drops a NaN operand; a NaN row reports +inf, which every max keeps,
  // so that non-finite data end the solve as diverged
  const double sum = e0 + e1 + e2;
  return sum == sum ? fmax(fabs(e0), fmax(fabs(e1), fabs(e2))) : INFINITY;
}

__device__ __forceinline__ void zero(double (&acc)[kSlots]) {
#pragma unroll
  for (int s = 0; s < kSlots; ++s) acc[s] = 0.0;
}

// loop of a workgroup over the row tiles (uniform across the workgroup)
#define TILE_LOOP(T) \
  for (int ti_ = tilesOf(T, (T).counts[0], gridDim.x).first, te_ = tilesOf(T, (T).counts[0], gridDim.x).end, \
           ts_ = tilesOf(T, (T).counts[0], gridDim.x).step;                                                 \
       ti_ < te_; ti_ += ts_)

__device__ __forceinline__ Tile rowTile(const Tables& T, int k) {
  return T.single ? Tile{0, k * kTC} : T.items[k];
}
__device__ __forceinline__ Tile bodyTile(const Tables& T, int k) {
  return T.single ? Tile{0, k * kTB} : T.bodies[k];
}

// partial-sum slots of a tile
#define TILE_PART(p, tile) ((p) + (long)((tile).start / kTC) * kSlots)

// XCD-contiguous tile order for the gathering kernels of a single slot.  The
// dispatcher hands workgroup b to XCD b % 8, and each XCD has an L2 of its
// own: workgroups b = x, x + 8, x + 16, ... take the tiles of the x-th eighth
// of [0, nt) in order, so that an XCD gathers from the bodies (and contacts)
// of one contiguous range, whose neighbours it mostly shares lines with,
// instead of every XCD fetching lines from all over the arrays.  Any grid of
// at least 8 workgroups covers every tile once (fewer: plain grid stride).
// Only the order of independent tiles changes, so results do not.
constexpr int kXcds = 8;
struct TileRange {
  int first, end, step;
};
__device__ __forceinline__ TileRange xcdTiles(int b, int nw, int nt) {
  if (nw < kXcds) return TileRange{b, nt, nw};
  const int x = b % kXcds, j = b / kXcds;
  const int q = nt / kXcds, r = nt % kXcds;
  const int start = x * q + min(x, r), cnt = q + (x < r ? 1 : 0);
  const int wx = nw / kXcds + (x < nw % kXcds ? 1 : 0);  // workgroups on XCD x
  return TileRange{start + j, start + cnt, wx};
}

// tiles of this workgroup among nt, by the schedule's order (XCD-contiguous
// or grid stride) over nw workgroups
__device__ __forceinline__ TileRange tilesOf(const Tables& T, int nt, int nw) {
  return T.xcd ? xcdTiles(blockIdx.x, nw, nt) : TileRange{(int)blockIdx.x, nt, nw};
}

// ---------------------------------------------------------------------------
// setup (APGD.cu:195-223)

// r to component-major; gamma_0, gammaHat_0 = 1, d = gamma_0 - gammaHat_0 (in
// ring[1]); J = F^T d; |d|^2.  The mixed solver first rounds the frames.
template <typename TM>
__global__ __launch_bounds__(kTC) void kSetup1(const Args<TM>* __restrict__ A, Tables T) {
  constexpr bool kMixed = sizeof(TM) < sizeof(double);
  TILE_LOOP(T) {
    const Tile tile = rowTile(T, ti_);
    const Args<TM>& a = A[tile.scene];
    const SolveDims& D = a.dims;
    const long items = (long)D.nc + D.nbil, n3 = 3L * D.nc;
    const long q = (long)tile.start + threadIdx.x;
    double acc[kSlots];
    zero(acc);
    if (q < items) {
      const bool bil = q >= D.nc;
      const int nr = bil ? 1 : 3;
      double x[3] = {0, 0, 0};
      for (int rr = 0; rr < nr; ++rr) {
        const long f = bil ? (q - D.nc) : (D.nbil + 3L * q + rr);  // flat (reference) row
        const long o = bil ? BROW(a, q - D.nc) : ROW(a, rr, q);
        gl(a.r)[o] = (TM)gl(a.rFlat)[f];
        // flat entries 0..3nc-1 start as (1,0,0) triples, without the
        // bilateral offset (APGD.cu:116-122); the rest keep their old values
        const double gv = (!D.warm && f < n3) ? ((f % 3 == 0) ? 1.0 : 0.0) : gl(a.gammaIO)[f];
        gl(a.ring[0])[o] = gv;
        gl(a.gHat)[o] = 1.0;
        x[rr] = 1.0 * gv + (-1.0) * 1.0;
        gl(a.ring[1])[o] = x[rr];
        acc[0] += x[rr] * x[rr];
      }
      if (!bil) {
        if (kMixed) {
#pragma unroll
          for (int k = 0; k < 6; ++k) gl(a.frame)[k * a.cap + q] = (TM)gl(a.frame64)[k * a.cap + q];
        }
        TM F[9];
        ldFrame(a, (int)q, D.geomMode, F);
        storeJ(a, (int)q, F, (TM)x[0], (TM)x[1], (TM)x[2]);
      }
    }
    writePartials<kTC>(acc, TILE_PART(a.part1[0], tile));
  }
}

// P2: tmp = M^-1 D^T x for non-fixed bodies; X holds the bilateral rows of x.
// which: 0 = ring[1] (d; also rounds M^-1 for the mixed solver), 1 = ring[0]
// (gamma_0), 2 = ring[nxt] (gamma_new of the pass whose state is st[sIdx]),
// 3 = gHat
template <typename TM>
__device__ __forceinline__ const DV_GLOBAL double* bodySource(const Args<TM>& a, int which, const PassView& S) {
  return gl(which == 0 ? a.ring[1] : which == 1 ? a.ring[0] : which == 2 ? a.ring[0] + S.nxt * a.ringStride : a.gHat);
}

template <typename TM>
__device__ __forceinline__ TM bodyMinv(const Args<TM>& a, int g, int which, bool own = true) {
  constexpr bool kMixed = sizeof(TM) < sizeof(double);
  if (kMixed && which == 0) {
    const TM mi = (TM)gl(a.minv64)[g];
    if (own) gl(a.minv)[g] = mi;
    return mi;
  }
  return gl(a.minv)[g];
}

// The Lipschitz test's term of one body in the pass whose view is S, from
// the body's D^T gamma_new (s) and M^-1: FP64, s . (M^-1 s), whose sum over
// the bodies is gamma_new' N gamma_new; mixed, d . (M^-1 d) with d = s - D^T y
// and D^T y = (1 + beta) D^T gamma_k - beta D^T gamma_{k-1} from the ring of
// per-body D^T gamma (gamma_k on a restart), whose sum is Delta' N Delta for
// Delta = gamma_new - y.  The mixed solver keeps s in ring slot nxt.
template <typename TM>
__device__ __forceinline__ double lipschitzTerm(const Args<TM>& a, int g, const PassView& S, TM mi, TM s0, TM s1,
                                                TM s2) {
  constexpr bool kMixed = sizeof(TM) < sizeof(double);
  if (!kMixed) return (double)s0 * (double)(mi * s0) + (double)s1 * (double)(mi * s1) + (double)s2 * (double)(mi * s2);
  const long nb = a.dims.nb;
  const V4<TM> c = ld4(a.sr, S.cur * nb + g), p = ld4(a.sr, S.prev * nb + g);
  V4<TM> sn;
  sn.x = s0; sn.y = s1; sn.z = s2; sn.w = 0;
  st4(a.sr, S.nxt * nb + g, sn);
  const bool mom = !(S.flags & kViewRestart);
  const double bp = 1.0 + S.beta, bm = -S.beta;
  const double d0 = (double)s0 - (mom ? bp * (double)c.x + bm * (double)p.x : (double)c.x);
  const double d1 = (double)s1 - (mom ? bp * (double)c.y + bm * (double)p.y : (double)c.y);
  const double d2 = (double)s2 - (mom ? bp * (double)c.z + bm * (double)p.z : (double)c.z);
  return (double)mi * (d0 * d0 + d1 * d1 + d2 * d2);
}

template <typename TM, int which>
__global__ __launch_bounds__(kTB) void kBody(const Args<TM>* __restrict__ A, Tables T, int sIdx) {
  constexpr bool kMixed = sizeof(TM) < sizeof(double);
  static_assert(kTB == kTP, "P2's extra workgroup reduces P1's partials like the decision");
  // the body workgroups (a single slot's pass has one more, the last)
  const int nw = (which == 2 && T.single) ? (int)gridDim.x - 1 : (int)gridDim.x;
  if ((int)blockIdx.x == nw) {
    // the pass's P1 partials (parity sIdx ^ 1), totalled while the other
    // workgroups gather, so that the decision reads eight numbers instead of
    // every tile's (after the solve has ended the totals are simply not used)
    const Args<TM>& a = A[0];
    const int par = sIdx ^ 1;
    double t1[kSlots];
    reduceSlots<kSlots, kSlots, 2, kP1Slots<TM>>(a.part1[par], 0, a.nTiles, t1);  // (two tiles' loads at a time: the same order)
    if (threadIdx.x < kSlots) gl(a.p1sum[par])[threadIdx.x] = t1[threadIdx.x];
    return;
  }
  auto tileBody = [&](const Tile tile, const PassView& S, const Args<TM>& a) {
    // no lane leaves early: lanes past the last body read body 0's entries
    // (unused), so that the index loads and M^-1 are all requested at once
    const int g = tile.start + threadIdx.x;
    const bool own = g < a.dims.nb;
    const int gs = own ? g : 0;
    const int k0 = gl(a.iStart)[gs], k1r = gl(a.iStart)[gs + 1];
    const int c0 = gl(a.cStart)[gs], c1r = gl(a.cStart)[gs + 1];
    const int b0 = gl(a.bilIncStart)[gs], b1r = gl(a.bilIncStart)[gs + 1];
    const TM mi = bodyMinv(a, gs, which, own);
    const bool heavy = k1r - k0 > kHeavyDegree;  // reduced below by a workgroup
    const bool sums = own && !heavy && mi != (TM)0;
    const int b1 = sums ? b1r : b0;
    // the body's incidences in ascending contact order, as two contiguous
    // runs: as body B, JB[k0 - c0, k1 - c1) (lower-numbered partners' rows),
    // then as body A, J[c0, c1) (its own row)
    const int jb0 = k0 - c0, jb1 = sums ? k1r - c1r : jb0;
    const int c1 = sums ? c1r : c0;
    const DV_GLOBAL double* X = bodySource(a, which, S);
    const long bil0 = 3L * a.cap;
    V4<TM> out;
    out.x = 0; out.y = 0; out.z = 0; out.w = 0;
    {
      TM s0 = 0, s1 = 0, s2 = 0;
      for (int k = b0; k < b1; ++k) {
        const int e = gl(a.bilInc)[k];
        const TM t = (TM)((e & 1) ? -X[bil0 + (e >> 3)] : X[bil0 + (e >> 3)]);
        const int comp = (e >> 1) & 3;
        if (comp == 0) s0 += t; else if (comp == 1) s1 += t; else s2 += t;
      }
      // runs in batches of 8 loads, then the sums in order, so the loads
      // overlap (slots past a run's end repeat its first entry, so that
      // every load is unconditional)
      for (int kb = jb0; kb < jb1; kb += 8) {
        V4<TM> j[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) j[u] = ld4(a.JB, kb + u < jb1 ? kb + u : kb);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          if (kb + u >= jb1) continue;
          s0 -= j[u].x; s1 -= j[u].y; s2 -= j[u].z;
        }
      }
      for (int kb = c0; kb < c1; kb += 8) {
        V4<TM> j[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) j[u] = ld4(a.J, kb + u < c1 ? kb + u : kb);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          if (kb + u >= c1) continue;
          s0 += j[u].x; s1 += j[u].y; s2 += j[u].z;
        }
      }
      out.x = mi * s0; out.y = mi * s1; out.z = mi * s2;
      if (which == 2) {
        // this body's Lipschitz term; the tile's sum goes to lq (bodies reduced
        // by a workgroup add theirs below)
        double term[1] = {own && !heavy ? lipschitzTerm(a, gs, S, mi, s0, s1, s2) : 0.0};
        __shared__ double tt[1];
        blockReduce<kTB, 1>(term, tt);
        if (threadIdx.x == 0) gl(a.lq[sIdx ^ 1])[tile.start / kTB] = tt[0];
      }
      if (which == 1 && kMixed && own && !heavy) {  // D^T gamma_0 into ring slot 0
        V4<TM> sv;
        sv.x = s0; sv.y = s1; sv.z = s2; sv.w = 0;
        st4(a.sr, g, sv);
      }
    }
    if (own && !heavy) st4(a.tmp, g, out);
  };
  if (T.single) {  // as in kP1
    const PassView S = loadView(T, 0, sIdx);
    const Args<TM>& a = A[0];
    const int nt = T.counts[1];
    if (!(which == 2 && viewDone(S))) {
      const TileRange tr = tilesOf(T, nt, nw);
      for (int ti = tr.first; ti < tr.end; ti += tr.step) tileBody(Tile{0, ti * kTB}, S, a);
    }
  } else {
    const TileRange tr = tilesOf(T, T.counts[1], nw);
    for (int ti = tr.first; ti < tr.end; ti += tr.step) {
      const Tile tile = bodyTile(T, ti);
      const PassView S = loadView(T, tile.scene, sIdx);
      if (which == 2 && viewDone(S)) continue;
      tileBody(tile, S, A[tile.scene]);
    }
  }
  for (int hq = blockIdx.x, nh = T.counts[2]; hq < nh; hq += nw) {
    const int2 ref = T.heavy[hq];
    const PassView S = loadView(T, ref.x, sIdx);
    if (which == 2 && viewDone(S)) continue;
    const Args<TM>& a = A[ref.x];
    const int g = ref.y;
    const TM mi = bodyMinv(a, g, which);
    // this body's slot among the Lipschitz partials, after the body tiles'
    double* lqSlot = which == 2 ? a.lq[sIdx ^ 1] + a.nBodyTiles + T.heavyIdx[hq] : nullptr;
    if (mi == (TM)0) {
      if (threadIdx.x == 0) {
        V4<TM> out;
        out.x = 0; out.y = 0; out.z = 0; out.w = 0;
        st4(a.tmp, g, out);
        if (which == 2) gl(lqSlot)[0] = 0.0;
      }
      continue;
    }
    const DV_GLOBAL double* X = bodySource(a, which, S);
    const long bil0 = 3L * a.cap;
    const int k0 = gl(a.iStart)[g], k1 = gl(a.iStart)[g + 1];
    const int c0 = gl(a.cStart)[g], c1 = gl(a.cStart)[g + 1];
    const int nd = (k1 - k0) - (c1 - c0), jb0 = k0 - c0;  // B-side incidences first
    double acc[kSlots];
    zero(acc);
    for (int k = threadIdx.x; k < k1 - k0; k += kTB) {
      const bool bSide = k < nd;
      const V4<TM> j = bSide ? ld4(a.JB, jb0 + k) : ld4(a.J, c0 + k - nd);
      const double sg = bSide ? -1.0 : 1.0;
      acc[0] += sg * (double)j.x;
      acc[1] += sg * (double)j.y;
      acc[2] += sg * (double)j.z;
    }
    if (threadIdx.x == 0) {
      for (int k = gl(a.bilIncStart)[g]; k < gl(a.bilIncStart)[g + 1]; ++k) {
        const int e = gl(a.bilInc)[k];
        acc[(e >> 1) & 3] += (e & 1) ? -X[bil0 + (e >> 3)] : X[bil0 + (e >> 3)];
      }
    }
    __shared__ double tot[kSlots];
    blockReduceR<kTB, 1, kSlots, 0x07u>(reinterpret_cast<double(&)[1][kSlots]>(acc), tot);
    if (threadIdx.x == 0) {
      V4<TM> out;
      out.x = mi * (TM)tot[0];
      out.y = mi * (TM)tot[1];
      out.z = mi * (TM)tot[2];
      out.w = 0;
      st4(a.tmp, g, out);
      if (which == 2) gl(lqSlot)[0] = lipschitzTerm(a, g, S, mi, (TM)tot[0], (TM)tot[1], (TM)tot[2]);
      if (which == 1 && kMixed) {
        V4<TM> sv;
        sv.x = (TM)tot[0]; sv.y = (TM)tot[1]; sv.z = (TM)tot[2]; sv.w = 0;
        st4(a.sr, g, sv);
      }
    }
  }
}

// |N d|^2 into part3[1]; J = F^T gamma_0
template <typename TM>
__global__ __launch_bounds__(kTC) void kSetup3(const Args<TM>* __restrict__ A, Tables T) {
  TILE_LOOP(T) {
    const Tile tile = rowTile(T, ti_);
    const Args<TM>& a = A[tile.scene];
    const SolveDims& D = a.dims;
    const long items = (long)D.nc + D.nbil;
    const long q = (long)tile.start + threadIdx.x;
    double acc[kSlots];
    zero(acc);
    if (q < items) {
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
        storeJ(a, c, F, (TM)gl(a.ring[0])[ROW(a, 0, c)], (TM)gl(a.ring[0])[ROW(a, 1, c)], (TM)gl(a.ring[0])[ROW(a, 2, c)]);
      }
    }
    writePartials<kTC>(acc, TILE_PART(a.part3[0], tile));
  }
}

// N gamma_0 into rn[0]
template <typename TM>
__global__ __launch_bounds__(kTC) void kSetup5(const Args<TM>* __restrict__ A, Tables T) {
  TILE_LOOP(T) {
    const Tile tile = rowTile(T, ti_);
    const Args<TM>& a = A[tile.scene];
    const SolveDims& D = a.dims;
    const long items = (long)D.nc + D.nbil;
    const long q = (long)tile.start + threadIdx.x;
    if (q >= items) continue;
    if (q >= D.nc) {
      gl(a.rn[0])[BROW(a, q - D.nc)] = bilRow(a, (int)(q - D.nc));
    } else {
      const int c = (int)q;
      TM F[9];
      ldFrame(a, c, D.geomMode, F);
      TM n0, n1, n2;
      contactRows(a, c, F, n0, n1, n2);
      gl(a.rn[0])[ROW(a, 0, c)] = n0; gl(a.rn[0])[ROW(a, 1, c)] = n1; gl(a.rn[0])[ROW(a, 2, c)] = n2;
    }
  }
}

// L_0 = |N d| / |d|, t_0 = 1/L_0 and the initial state (APGD.cu:208-223),
// written to both parity slots; one wavefront per scene
template <typename TM>
__global__ __launch_bounds__(kTD) void kSetupDecide(const Args<TM>* __restrict__ A, Tables T) {
  const int slot = T.slots[blockIdx.x];
  const Args<TM>& a = A[slot];
  double tot[2 * kSlots];
  reduceAll(a.part1[0], a.part3[0], a.nTiles, tot);
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
    S.it = 0; S.bt = 0; S.backtracks = 0; S.restart = 1;
    S.cur = 0; S.prev = 0; S.nxt = 1; S.hatSrc = -1;  // -1: gHat (ones)
    S.pendingRes = 0;  // gamma_0 is not an accepted iterate (APGD.cu:228-230)
    S.fresh = 0;       // N gamma_0 is in rn[0] (kSetup5)
    S.done = a.dims.maxIt <= 0 ? 1 : 0;
    S.converged = 0;
    S.passes = 0;
    STATE(T, slot, 0) = S;
    STATE(T, slot, 1) = S;
    VIEW(T, slot, 0) = viewOf(S);
    VIEW(T, slot, 1) = viewOf(S);
    T.out[slot] = S;
  }
}

// ---------------------------------------------------------------------------
// the decision of a pass: Lipschitz test, backtracking or acceptance
// (APGD.cu:236-332), from the pass's fixed-order partial sums tot[0..15]

template <bool kMixed>
__device__ void decideStep(ApgdState& S, const double* tot, const SolveDims& D) {
  const double* s1 = tot;           // P1 slots
  const double* s3 = tot + kSlots;  // slot 6: P2's Lipschitz term
  S.passes++;
  // getResidual of the iterate the previous decision accepted (now in slot
  // cur), evaluated by this pass's P1, then gammaHat and the stopping test
  // (APGD.cu:276-289).  Taking it one pass late spares a pass the reload of the iterate and r;
  // the decisions and the result are those of the reference's order.
  if (S.pendingRes) {
    S.pendingRes = 0;
    const double res = s1[kSlots - 1];
    if (!(res < INFINITY)) {  // non-finite data: the solve ends as diverged
      S.best = res;
      S.done = 1;
      return;
    }
    if (res < S.best) {  // gammaHat := that iterate, kept in its ring slot
      S.best = res;
      S.hatSrc = S.cur;
    }
    if (S.best < D.tol) {
      S.done = 1;
      S.converged = 1;
      S.it -= 1;  // the reference stops inside the iteration it counts
      return;
    }
    if (S.it >= D.maxIt) {
      S.done = 1;
      return;
    }
  }
  const double L = S.L;
  bool backtrack;
  if (kMixed) {
    // P2's body sums give Delta' N Delta = |M^-1/2 D^T Delta|^2 directly
    backtrack = s3[6] >= L * s1[4];
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
    S.fresh = 0;  // this pass's P1 has formed and stored N of slot cur
    return;
  }
  const double theta = S.theta;
  double thetaNew = (-(theta * theta) + theta * sqrt(theta * theta + 4.0)) / 2.0;
  S.beta = theta * (1.0 - theta) / (theta * theta + thetaNew);
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
  S.pendingRes = 1;
  S.fresh = 1;  // gamma_new's N is formed by the next P1, from the tmp of this pass's P2
}

// ---------------------------------------------------------------------------
// one pass (parity par): the state of the pass is st[par ^ 1]

// Frame of a contact from its loaded components: FP64 stores n (u, v are
// rebuilt bit-identically, System.cu:753-762), FP32 stores n and u.
__device__ __forceinline__ void frameFrom(const double (&c)[6], int mode, double (&F)[9]) {
  const d3 n = mk3(c[0], c[1], c[2]);
  d3 u, v;
  contactFrame(n, mode, &u, &v);
  F[0] = n.x; F[1] = n.y; F[2] = n.z;
  F[3] = u.x; F[4] = u.y; F[5] = u.z;
  F[6] = v.x; F[7] = v.y; F[8] = v.z;
}
__device__ __forceinline__ void frameFrom(const float (&c)[6], int, float (&F)[9]) {
#pragma unroll
  for (int k = 0; k < 6; ++k) F[k] = c[k];
  F[6] = F[1] * F[5] - F[2] * F[4];  // v = n x u
  F[7] = F[2] * F[3] - F[0] * F[5];
  F[8] = F[0] * F[4] - F[1] * F[3];
}

// The contact kernels take R rows (contacts) per thread: a workgroup of
// kTC / R threads covers one tile of kTC rows, and every thread issues the
// loads of all its rows before computing, for more memory-level parallelism
// per wavefront.  The per-tile partial sums keep their layout.
//
// The mixed P1 with one row per thread is held to 80 VGPRs (six wavefronts
// per SIMD), so that a CU keeps three of its 512-thread workgroups in flight
// rather than two: past two tiles per CU (262,144 rows on 256 CUs), the CUs
// given a third tile then run it beside the other two instead of after one
// of them (P1 11% faster at 263k contacts, 2% slower at 256k, where no CU
// takes a third tile).  The FP64 instance, which holds twice the registers
// per value, needs 94 VGPRs (five wavefronts); it is held to no fewer than
// four, so that it does not spill.
template <typename TM, int R>
constexpr int kP1MinWaves = R == 1 ? (sizeof(TM) < sizeof(double) ? 6 : 4) : 1;

// The decision of the previous pass (parity par ^ 1) of a single slot, taken
// by every workgroup of the next pass's P1 instead of by a kernel of its own:
// the same state, totals and summation order as kDecide give the same new
// state, which workgroup 0 stores, and the pass view follows from it without
// a reload.  P1 reads state par and writes state par ^ 1 and only its own
// partial sums, so a repeated launch reproduces the same results.
template <typename TM>
__device__ PassView decidePrevious(const Args<TM>& a, const Tables& T, int par) {
  constexpr bool kMixed = sizeof(TM) < sizeof(double);
  ApgdState S = STATE(T, 0, par);
  double tot[2 * kSlots];
  passTotals(a.p1sum[par ^ 1], a.lq[par ^ 1], a.nBodyTiles + gl(a.counters)[2], tot);
  if (!S.done) decideStep<kMixed>(S, tot, a.dims);
  const PassView v = viewOf(S);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    STATE(T, 0, par ^ 1) = S;
    VIEW(T, 0, par ^ 1) = v;
    T.out[0] = S;
  }
  return v;
}

// kFold: a single slot's P1 takes the previous pass's decision first.
template <typename TM, int R, bool kFold>
__global__ __launch_bounds__(kTC / R, (kP1MinWaves<TM, R>)) void kP1(const Args<TM>* __restrict__ A, Tables T, int par) {
  constexpr bool kMixed = sizeof(TM) < sizeof(double);
  constexpr int NT = kTC / R;
  constexpr int NF = kMixed ? 6 : 3;  // stored frame components
  // the contacts' body ids and both bodies' tmp (M^-1 D^T of the iterate the
  // previous pass formed), from which the rows of N gamma_k are formed when
  // the previous decision accepted that iterate: a single slot loads them
  // before it takes the decision, so that they arrive while it decides
  // (after a backtrack the gathered tmp is not used)
  // the loads that do not depend on the pass view (r, the frame, the B-side
  // slot), issued before the view (and a single slot's decision) arrives
  struct Pre {
    TM rv[R][3], fc[R][6];
    int pb[R];
  };
  auto loadPre = [&](const Tile tile, const Args<TM>& a, Pre& pr) {
    const SolveDims& D = a.dims;
    const long items = (long)D.nc + D.nbil;
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const long q = (long)tile.start + threadIdx.x + u * NT;
      const bool valid = q < items, bil = q >= D.nc;
      const long qf = (valid && !bil) ? q : 0;
#pragma unroll
      for (int rr = 0; rr < 3; ++rr) {
        const bool on = valid && (rr == 0 || !bil);
        const long f = on ? (bil ? BROW(a, q - D.nc) : ROW(a, rr, q)) : 0;
        pr.rv[u][rr] = gl(a.r)[f];
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) pr.fc[u][k] = k < NF ? gl(a.frame)[k * a.cap + qf] : (TM)0;
      pr.pb[u] = gl(a.posB)[qf];
    }
  };
  auto loadIds = [&](const Tile tile, const Args<TM>& a, int (&ia)[R], int (&ib)[R], V4<TM> (&tA)[R], V4<TM> (&tB)[R]) {
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const long q = (long)tile.start + threadIdx.x + u * NT;
      const long qf = q < a.dims.nc ? q : 0;
      ia[u] = gl(a.cA)[qf];
      ib[u] = gl(a.cB)[qf];
    }
    // (rows that are not a contact gather body 0: the id arrays hold at
    // least one entry, but maybe no contact)
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const bool own = (long)tile.start + threadIdx.x + u * NT < a.dims.nc;
      tA[u] = ld4(a.tmp, own ? ia[u] : 0);
      tB[u] = ld4(a.tmp, own ? ib[u] : 0);
    }
  };
  auto tileBody = [&](const Tile tile, const PassView& S, const Args<TM>& a, const int (&ia)[R], const int (&ib)[R],
                      const V4<TM> (&tA)[R], const V4<TM> (&tB)[R], const Pre& pr) {
    const SolveDims& D = a.dims;
    const long items = (long)D.nc + D.nbil;
    const DV_GLOBAL double* Gk = gl(a.ring[0] + S.cur * a.ringStride);
    const DV_GLOBAL double* Gkm1 = gl(a.ring[0] + S.prev * a.ringStride);
    const DV_GLOBAL TM* Nk = gl(a.rn[0] + S.cur * a.ringStride);
    const DV_GLOBAL TM* Nkm1 = gl(a.rn[0] + S.prev * a.ringStride);
    DV_GLOBAL double* Gn = gl(a.ring[0] + S.nxt * a.ringStride);
    const bool mom = !(S.flags & kViewRestart);  // momentum term present
    // loads of all rows of the thread, issued before any of them is used:
    // rows the thread does not own read row 0 instead (always present), so
    // that no load is conditional and no select waits on one; their values
    // are never used
    // N gamma_k: formed here from the tmp of the previous pass's P2 when the
    // previous decision accepted gamma_k (fresh), and stored for the passes
    // after; read back otherwise (the flag is the same in every workgroup)
    const bool fresh = (S.flags & kViewFresh) != 0;
    double gk[R][3], gm[R][3];
    TM nk[R][3], nm[R][3];
    const auto& rv = pr.rv;
    const auto& fc = pr.fc;
    const auto& pb = pr.pb;
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const long q = (long)tile.start + threadIdx.x + u * NT;
      const bool valid = q < items, bil = q >= D.nc;
#pragma unroll
      for (int rr = 0; rr < 3; ++rr) {
        const bool on = valid && (rr == 0 || !bil);
        const long f = on ? (bil ? BROW(a, q - D.nc) : ROW(a, rr, q)) : 0;
        gk[u][rr] = Gk[f];
        gm[u][rr] = Gkm1[f];  // read even after a restart (unused then)
        nm[u][rr] = Nkm1[f];
        nk[u][rr] = (TM)0;
      }
    }
    if (fresh) {
      DV_GLOBAL TM* Nw = gl(a.rn[0] + S.cur * a.ringStride);
#pragma unroll
      for (int u = 0; u < R; ++u) {
        const long q = (long)tile.start + threadIdx.x + u * NT;
        if (q >= items) continue;
        if (q >= D.nc) {
          nk[u][0] = bilRow(a, (int)(q - D.nc));
          Nw[BROW(a, q - D.nc)] = nk[u][0];
        } else {
          TM F[9];
          frameFrom(fc[u], D.geomMode, F);
          rowsOfTmp(F, tA[u], tB[u], nk[u][0], nk[u][1], nk[u][2]);
#pragma unroll
          for (int rr = 0; rr < 3; ++rr) Nw[ROW(a, rr, q)] = nk[u][rr];
        }
      }
    } else {
#pragma unroll
      for (int u = 0; u < R; ++u) {
        const long q = (long)tile.start + threadIdx.x + u * NT;
        const bool valid = q < items, bil = q >= D.nc;
#pragma unroll
        for (int rr = 0; rr < 3; ++rr) {
          const bool on = valid && (rr == 0 || !bil);
          nk[u][rr] = Nk[on ? (bil ? BROW(a, q - D.nc) : ROW(a, rr, q)) : 0];
        }
      }
    }
    const double bp = 1.0 + S.beta, bm = -S.beta;
    const double t = S.t;
    double acc[R][kSlots];  // per row of the thread: see blockReduceR
#pragma unroll
    for (int u = 0; u < R; ++u) zero(acc[u]);
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const long q = (long)tile.start + threadIdx.x + u * NT;
      if (q >= items) continue;
      const bool bil = q >= D.nc;
      const int nr = bil ? 1 : 3;
      double y[3] = {0, 0, 0}, ny[3] = {0, 0, 0}, gg[3] = {0, 0, 0}, gn[3] = {0, 0, 0};
#pragma unroll
      for (int rr = 0; rr < 3; ++rr) {
        if (rr < nr) {
          if (mom) {
            y[rr] = bp * gk[u][rr] + bm * gm[u][rr];
            ny[rr] = bp * (double)nk[u][rr] + bm * (double)nm[u][rr];
          } else {
            y[rr] = gk[u][rr];
            ny[rr] = (double)nk[u][rr];
          }
          gg[rr] = 1.0 * ny[rr] + 1.0 * (double)rv[u][rr];
          gn[rr] = 1.0 * y[rr] + (-t) * gg[rr];
        }
      }
      // getResidual (APGD.cu:90-102) of gamma_k, with N gamma_k and r
      // already loaded: gamma_k is the iterate the previous decision
      // accepted, whose residual that decision left to this pass
      {
        const double w0 = (double)nk[u][0] + (double)rv[u][0];
        if (bil) {
          double e;
          if (D.nc) {
            const double z = 1.0 * gk[u][0] + (-D.gdiff) * w0;
            const double ig = D.igdiff;
            e = fabs(ig * gk[u][0] + (-ig) * z);
          } else {
            e = fabs(w0);
          }
          acc[u][7] = fmax(acc[u][7], e == e ? e : INFINITY);
        } else {
          const double w1 = (double)nk[u][1] + (double)rv[u][1];
          const double w2 = (double)nk[u][2] + (double)rv[u][2];
          acc[u][7] = fmax(acc[u][7], residualContact(gk[u][0], gk[u][1], gk[u][2], w0, w1, w2, D.mu, D.gdiff, D.igdiff));
        }
      }
      if (!bil) projectCone<double>(gn[0], gn[1], gn[2], D.mu);
#pragma unroll
      for (int rr = 0; rr < 3; ++rr) {
        if (rr < nr) {
          const long f = bil ? BROW(a, q - D.nc) : ROW(a, rr, q);
          Gn[f] = gn[rr];
          const double dy = gn[rr] - y[rr];
          if (!kMixed) {
            acc[u][0] += gn[rr] * (double)rv[u][rr];
            acc[u][1] += y[rr] * ny[rr];
            acc[u][2] += y[rr] * (double)rv[u][rr];
            acc[u][3] += gg[rr] * dy;
          }
          acc[u][4] += dy * dy;
          acc[u][5] += gg[rr] * (gn[rr] - gk[u][rr]);
        }
      }
      if (!bil) {
        TM F[9];
        frameFrom(fc[u], D.geomMode, F);
        storeJ(a, (int)q, pb[u], F, (TM)gn[0], (TM)gn[1], (TM)gn[2]);
      }
    }
    // the slots this build adds to (the others are stored as zeros): the
    // mixed solver's Lipschitz test takes only |Delta|^2 from P1
    writePartialsR<NT, R, kP1Slots<TM>>(acc, TILE_PART(a.part1[par], tile));
  };
  if (T.single) {
    // one slot: the grid has a workgroup for every tile the slot's capacity
    // can hold (the captured graph outlives changes of the row count), so
    // each takes one tile and nothing stays live across a tile loop (which
    // held P1 above 80 VGPRs).  A workgroup whose tile lies past the row
    // count leaves once the count has arrived, before the decision's
    // reductions (its partial sums are never read); workgroup 0 always
    // stays, to store the decided state.  The count comes with the table
    // entry, which the id loads wait for anyway.
    const Args<TM>& a = A[0];
    const Tile tile{0, (int)blockIdx.x * kTC};
    int ia[R], ib[R];
    V4<TM> tA[R], tB[R];
    loadIds(tile, a, ia, ib, tA, tB);
    if (blockIdx.x > 0 && (long)tile.start >= (long)a.dims.nc + a.dims.nbil) return;
    Pre pr;
    loadPre(tile, a, pr);
    const PassView S = kFold ? decidePrevious(a, T, par) : loadView(T, 0, par ^ 1);
    if (viewDone(S)) return;
    tileBody(tile, S, a, ia, ib, tA, tB, pr);
    return;
  }
  TILE_LOOP(T) {
    const Tile tile = rowTile(T, ti_);
    const PassView S = loadView(T, tile.scene, par ^ 1);
    if (viewDone(S)) continue;
    int ia[R], ib[R];
    V4<TM> tA[R], tB[R];
    loadIds(tile, A[tile.scene], ia, ib, tA, tB);
    Pre pr;
    loadPre(tile, A[tile.scene], pr);
    tileBody(tile, S, A[tile.scene], ia, ib, tA, tB, pr);
  }
}

// (N x)_c = F_c (tmp_A - tmp_B) and the bilateral rows of x whose
// M^-1 D^T x is in tmp, into N of ring slot nxt (the view of parity par ^ 1),
// one row tile per workgroup of a single slot: the rows a pass's P1 forms
// for a fresh iterate, as a kernel of their own for dv_pass_product.
template <typename TM>
__global__ __launch_bounds__(kTC) void kRowsN(const Args<TM>* __restrict__ A, Tables T, int par) {
  const PassView S = loadView(T, 0, par ^ 1);
  const Args<TM>& a = A[0];
  const SolveDims& D = a.dims;
  const long items = (long)D.nc + D.nbil;
  const long q = (long)blockIdx.x * kTC + threadIdx.x;
  if (q >= items) return;
  DV_GLOBAL TM* Nn = gl(a.rn[0] + S.nxt * a.ringStride);
  if (q >= D.nc) {
    Nn[BROW(a, q - D.nc)] = bilRow(a, (int)(q - D.nc));
    return;
  }
  TM F[9];
  ldFrame(a, (int)q, D.geomMode, F);
  TM n0, n1, n2;
  rowsOfTmp(F, ld4(a.tmp, gl(a.cA)[q]), ld4(a.tmp, gl(a.cB)[q]), n0, n1, n2);
  Nn[ROW(a, 0, q)] = n0; Nn[ROW(a, 1, q)] = n1; Nn[ROW(a, 2, q)] = n2;
}

// Decision of the pass of parity par for every slot (one wavefront each):
// state st[par ^ 1] -> st[par], recorded in T.out for the host.  A slot that
// is done copies its state forward.
template <typename TM>
__global__ __launch_bounds__(kTP) void kDecide(const Args<TM>* __restrict__ A, Tables T, int par) {
  constexpr bool kMixed = sizeof(TM) < sizeof(double);
  if (T.single) {
    // one slot, always active: the table entry and the state are requested
    // together, and the partials are reduced without waiting for the done
    // flag (after the solve has ended the totals are simply not used)
    const Args<TM>& a = A[0];
    ApgdState S = STATE(T, 0, par ^ 1);
    double tot[2 * kSlots];
    passTotals(a.p1sum[par], a.lq[par], a.nBodyTiles + gl(a.counters)[2], tot);
    if (!S.done) decideStep<kMixed>(S, tot, a.dims);
    if (threadIdx.x == 0) {
      STATE(T, 0, par) = S;
      VIEW(T, 0, par) = viewOf(S);
      T.out[0] = S;
    }
    return;
  }
  const Args<TM>& a = A[blockIdx.x];  // one workgroup per slot
  if (!a.active) return;
  ApgdState S = STATE(T, blockIdx.x, par ^ 1);
  if (!S.done) {
    double tot[2 * kSlots];
    reducePass<TM>(a.part1[par], a.nTiles, a.lq[par], a.nBodyTiles + gl(a.counters)[2], tot);
    decideStep<kMixed>(S, tot, a.dims);
  }
  if (threadIdx.x == 0) {
    STATE(T, blockIdx.x, par) = S;
    VIEW(T, blockIdx.x, par) = viewOf(S);
    T.out[blockIdx.x] = S;
  }
}

// ---------------------------------------------------------------------------
// gamma := gammaHat (APGD.cu:358) and objectiveCCP (APGD.cu:353-354); the
// final state is st[fin]

template <typename TM>
__global__ __launch_bounds__(kTC) void kFin1(const Args<TM>* __restrict__ A, Tables T, int fin) {
  TILE_LOOP(T) {
    const Tile tile = rowTile(T, ti_);
    const Args<TM>& a = A[tile.scene];
    const SolveDims& D = a.dims;
    const long items = (long)D.nc + D.nbil;
    const long q = (long)tile.start + threadIdx.x;
    const int hatSrc = STATE(T, tile.scene, fin).hatSrc;
    const DV_GLOBAL double* hatV = gl(hatSrc >= 0 ? a.ring[hatSrc] : a.gHat);
    double acc[kSlots];
    zero(acc);
    if (q < items) {
      const bool bil = q >= D.nc;
      const int nr = bil ? 1 : 3;
      double x[3] = {0, 0, 0};
      for (int rr = 0; rr < nr; ++rr) {
        const long o = bil ? BROW(a, q - D.nc) : ROW(a, rr, q);
        const long f = bil ? (q - D.nc) : (D.nbil + 3L * q + rr);
        x[rr] = hatV[o];
        gl(a.gHat)[o] = x[rr];
        gl(a.gammaIO)[f] = x[rr];
        acc[0] += x[rr] * (double)gl(a.r)[o];
      }
      if (!bil) {
        TM F[9];
        ldFrame(a, (int)q, D.geomMode, F);
        storeJ(a, (int)q, F, (TM)x[0], (TM)x[1], (TM)x[2]);
      }
    }
    writePartials<kTC>(acc, TILE_PART(a.part1[0], tile));
  }
}

template <typename TM>
__global__ __launch_bounds__(kTC) void kFin3(const Args<TM>* __restrict__ A, Tables T) {
  TILE_LOOP(T) {
    const Tile tile = rowTile(T, ti_);
    const Args<TM>& a = A[tile.scene];
    const SolveDims& D = a.dims;
    const long items = (long)D.nc + D.nbil;
    const long q = (long)tile.start + threadIdx.x;
    double acc[kSlots];
    zero(acc);
    if (q < items) {
      if (q >= D.nc) {
        acc[1] += gl(a.gHat)[BROW(a, q - D.nc)] * (double)bilRow(a, (int)(q - D.nc));
      } else {
        const int c = (int)q;
        TM F[9];
        ldFrame(a, c, D.geomMode, F);
        TM n0, n1, n2;
        contactRows(a, c, F, n0, n1, n2);
        acc[1] += gl(a.gHat)[ROW(a, 0, c)] * (double)n0 + gl(a.gHat)[ROW(a, 1, c)] * (double)n1 +
                  gl(a.gHat)[ROW(a, 2, c)] * (double)n2;
      }
    }
    writePartials<kTC>(acc, TILE_PART(a.part3[0], tile));
  }
}

template <typename TM>
__global__ __launch_bounds__(kTD) void kFinDecide(const Args<TM>* __restrict__ A, Tables T, int fin) {
  const int slot = T.slots[blockIdx.x];
  const Args<TM>& a = A[slot];
  double tot[2 * kSlots];
  reduceAll(a.part1[0], a.part3[0], a.nTiles, tot);
  if (threadIdx.x == 0) {
    ApgdState S = STATE(T, slot, fin);
    S.objective = 0.5 * tot[kSlots + 1] + tot[0];
    STATE(T, slot, fin) = S;
    T.out[slot] = S;
  }
}

// (slot, body) list of the heavy bodies of the slots of a set, in slot-list
// order; one workgroup, slots in blocks of 1024 with a workgroup-wide prefix
// sum
template <typename TM>
__global__ __launch_bounds__(1024) void kHeavyTable(const Args<TM>* __restrict__ A, const int* slots, int nSlots,
                                                     int2* refs, int* refIdx, int* counts) {
  __shared__ int sc[1024];
  int base = 0;
  for (int s0 = 0; s0 < nSlots; s0 += 1024) {
    const int i = s0 + (int)threadIdx.x;
    const int s = i < nSlots ? slots[i] : 0;
    const int n = i < nSlots ? A[s].counters[2] : 0;
    sc[threadIdx.x] = n;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {  // inclusive scan
      const int v = (int)threadIdx.x >= d ? sc[threadIdx.x - d] : 0;
      __syncthreads();
      sc[threadIdx.x] += v;
      __syncthreads();
    }
    const int off = base + sc[threadIdx.x] - n;
    for (int h = 0; h < n; ++h) {
      refs[off + h] = make_int2(s, A[s].heavy[h]);
      refIdx[off + h] = h;  // its place in the slot's list, where P2 puts its Lipschitz term
    }
    base += sc[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[2] = base;
}

// ---------------------------------------------------------------------------
// host side

template <typename TM>
struct Plan {
  const Args<TM>* A;
  Tables T;
  int gc, gb, nSlots, rpt;  // rpt: rows per thread of the contact kernels
};

template <typename TM>
void launchP1(const Plan<TM>& P, hipStream_t st, int par, bool fold) {
  if (P.rpt == 2 && fold) hipLaunchKernelGGL((kP1<TM, 2, true>), dim3(P.gc), dim3(kTC / 2), 0, st, P.A, P.T, par);
  else if (P.rpt == 2) hipLaunchKernelGGL((kP1<TM, 2, false>), dim3(P.gc), dim3(kTC / 2), 0, st, P.A, P.T, par);
  else if (fold) hipLaunchKernelGGL((kP1<TM, 1, true>), dim3(P.gc), dim3(kTC), 0, st, P.A, P.T, par);
  else hipLaunchKernelGGL((kP1<TM, 1, false>), dim3(P.gc), dim3(kTC), 0, st, P.A, P.T, par);
}


constexpr int kProfBurst = 16;  // timed back-to-back launches of the profiled kernel

template <typename TM>
void launchPass(const Plan<TM>& P, hipStream_t st, int kern, int par, bool fold) {
  if (kern == 1) launchP1<TM>(P, st, par, fold);
  // a single slot's P2 has one more workgroup, which totals P1's partials
  if (kern == 2)
    hipLaunchKernelGGL((kBody<TM, 2>), dim3(P.gb + (P.T.single ? 1 : 0)), dim3(kTB), 0, st, P.A, P.T, par ^ 1);
  if (kern == 4) hipLaunchKernelGGL(kDecide<TM>, dim3(P.nSlots), dim3(kTP), 0, st, P.A, P.T, par);
}

// kProfBurst launches of pass kernel `kern` (P1 or P2) between ev[0] and
// ev[1], as a captured graph like the passes themselves, so that the average
// is the kernel's duration in a graph rather than that plus a stream
// launch's latency.  The graph is kept for the next solve of the same plan.
template <typename TM>
void timedBurst(const Plan<TM>& P, hipStream_t st, int kern, int par, bool fold, hipEvent_t* ev,
                GraphCache& cache) {
  struct Key {
    Plan<TM> P;
    int kern, par, fold;
  } key;
  std::memset(&key, 0, sizeof(key));
  key.P = P;
  key.kern = kern;
  key.par = par;
  key.fold = fold ? 1 : 0;
  if (!cache.matches(&key, sizeof(key))) {
    cache.reset();
    hipGraph_t graph;
    DV_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int r = 0; r < kProfBurst; ++r) launchPass<TM>(P, st, kern, par, fold);
    DV_HIP(hipStreamEndCapture(st, &graph));
    cache.instantiate(graph, "hipGraphInstantiate(timed burst)");
    cache.setKey(&key, sizeof(key));
    DV_HIP(hipGraphUpload(cache.exec, st));  // outside the timed region
  }
  DV_HIP(hipEventRecord(ev[0], st));
  cache.launch(st);
  DV_HIP(hipEventRecord(ev[1], st));
}

// One chunk of n (even) passes, pass k with parity k & 1.  A single slot's
// decision is taken by the next pass's P1, so its chunk launches the decide
// kernel once, after the last pass (the host reads the states after a
// chunk), and the first pass's P1 reads the view that kernel left.  With `ev`
// (stream launches only: events recorded inside a captured graph cannot be
// timed), kernel `prof` runs as a timed burst (P1 and P2 read nothing they
// write, apart from the N that a P1 forms for a fresh iterate and every
// repeat forms the same, so the repeats reproduce the same outputs) in the chunk's first
// pass, or for a P1 that decides, its second (whose decision has this
// chunk's partial sums); the decide kernel, which is not repeatable, is
// timed alone where it runs.
template <typename TM>
void enqueuePasses(const Plan<TM>& P, hipStream_t st, int n, hipEvent_t* ev, int prof, GraphCache* burst) {
  const bool single = P.T.single != 0;
  for (int k = 0; k < n; ++k) {
    const int par = k & 1;
    const bool fold = single && k > 0;
    for (int kern = 1; kern <= 4; ++kern) {
      if (kern == 3) continue;  // (a pass's contact rows are formed inside the next P1)
      if (kern == 4 && single && k != n - 1) continue;
      const int kt = !single ? 0 : kern == 1 ? 1 : kern == 4 ? n - 1 : 0;
      const bool timed = ev && prof == kern && k == kt;
      if (timed && kern != 4) {
        timedBurst<TM>(P, st, kern, par, fold, ev, *burst);
        continue;
      }
      if (timed) DV_HIP(hipEventRecord(ev[0], st));
      launchPass<TM>(P, st, kern, par, fold);
      if (timed) DV_HIP(hipEventRecord(ev[1], st));
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// The schedule tuner.  No schedule changes a result (see Schedule), so
// candidate schedules are tried on successive chunks of live solves, each
// chunk timed with events, and the solve goes on as it would under any of
// them.  Once every candidate of a class has a measurement, the fastest per
// pass and per active row becomes that class's schedule, kept for the
// process and, per device, appended to a text file ($DVIDE_SCHEDULE_FILE,
// else $HOME/.cache/dvide/schedules-<arch>-<CUs>.txt), from which later
// processes take it without measuring.  A class is the device, the build,
// one slot or several, and log2 of the rows solved.  DVIDE_TUNE=0 keeps the
// first candidate of each list.

namespace {

const Schedule kOneSlot[] = {{1, 1, 0, 4}, {1, 0, 0, 4}, {2, 1, 0, 4}, {2, 0, 0, 4}};
const Schedule kManySlots[] = {{1, 1, 32, 32}, {1, 0, 32, 32}, {2, 1, 32, 32},
                               {1, 1, 16, 16}, {1, 1, 64, 32}, {2, 1, 64, 32}};
constexpr int kOneSlotN = (int)(sizeof(kOneSlot) / sizeof(kOneSlot[0]));
constexpr int kManySlotsN = (int)(sizeof(kManySlots) / sizeof(kManySlots[0]));

struct TuneClass {
  std::vector<double> score;  // per candidate: us per pass per million active rows (< 0: not measured)
  int best = -1;              // the chosen candidate (-1: still measuring)
  Schedule sched;             // its schedule (also when read from the file)
};

struct Tuner {
  std::mutex m;
  std::map<std::string, TuneClass> classes;
  std::map<int, std::string> tags;  // device -> "<arch>-<CUs>"
  std::map<std::string, bool> loaded;
};

Tuner& tuner() {
  static Tuner t;
  return t;
}

bool tuningEnabled() {
  const char* e = getenv("DVIDE_TUNE");
  return !(e && e[0] == '0');
}

std::string scheduleFile(const std::string& tag) {
  const char* f = getenv("DVIDE_SCHEDULE_FILE");
  if (f && f[0]) return f;
  const char* home = getenv("HOME");
  if (!home || !home[0]) return "";
  return std::string(home) + "/.cache/dvide/schedules-" + tag + ".txt";
}

// caller holds the tuner's lock
std::string deviceTag(Tuner& T) {
  int dev = 0;
  DV_HIP(hipGetDevice(&dev));
  auto it = T.tags.find(dev);
  if (it != T.tags.end()) return it->second;
  hipDeviceProp_t p;
  DV_HIP(hipGetDeviceProperties(&p, dev));
  std::string arch = p.gcnArchName;
  arch = arch.substr(0, arch.find(':'));
  const std::string tag = arch + "-" + std::to_string(p.multiProcessorCount);
  T.tags[dev] = tag;
  if (!T.loaded[tag]) {
    T.loaded[tag] = true;
    const std::string path = scheduleFile(tag);
    if (FILE* f = path.empty() ? nullptr : fopen(path.c_str(), "r")) {
      char key[256];
      Schedule sc;
      double score;
      while (fscanf(f, "%255s %d %d %d %d %lf", key, &sc.rpt, &sc.xcd, &sc.wgc, &sc.wgb, &score) == 6) {
        if (sc.rpt < 1 || sc.rpt > 2 || sc.wgc < 0 || sc.wgb < 1) continue;
        TuneClass& c = T.classes[key];
        c.best = 0;  // decided
        c.sched = sc;
      }
      fclose(f);
    }
  }
  return tag;
}

void persist(const std::string& tag, const std::string& key, const Schedule& sc, double score) {
  const std::string path = scheduleFile(tag);
  if (path.empty()) return;
  if (!getenv("DVIDE_SCHEDULE_FILE")) {  // create $HOME/.cache/dvide
    const std::string dir = path.substr(0, path.rfind('/'));
    (void)mkdir(dir.substr(0, dir.rfind('/')).c_str(), 0755);
    (void)mkdir(dir.c_str(), 0755);
  }
  if (FILE* f = fopen(path.c_str(), "a")) {
    fprintf(f, "%s %d %d %d %d %.6g\n", key.c_str(), sc.rpt, sc.xcd, sc.wgc, sc.wgb, score);
    fclose(f);
  }
}

// The schedule of a solve: fixed rows per thread, the class's decided
// schedule, or (returned class, still measuring) the first candidate.
TuneClass* chooseSchedule(Engine& E, long rows, bool allowTuning) {
  const bool single = E.nSlots == 1;
  const Schedule* cands = single ? kOneSlot : kManySlots;
  const int n = single ? kOneSlotN : kManySlotsN;
  E.sched = cands[0];
  if (E.forcedRpt) {
    E.sched.rpt = E.forcedRpt;
    return nullptr;
  }
  if (!tuningEnabled()) return nullptr;  // the first candidate, whatever the schedule file holds
  Tuner& T = tuner();
  std::lock_guard<std::mutex> lock(T.m);
  int bucket = 0;
  while ((2L << bucket) <= rows) ++bucket;
  const std::string key = deviceTag(T) + ":" + (E.mixed ? "mixed" : "fp64") + ":" + (single ? "one" : "many") +
                          ":" + std::to_string(bucket);
  TuneClass& c = T.classes[key];
  if (c.best >= 0) {
    E.sched = c.sched;
    return nullptr;
  }
  if (!allowTuning) return nullptr;
  if ((int)c.score.size() != n) c.score.assign(n, -1.0);
  E.tuneKey = key;
  return &c;
}

// the next candidate to measure (not the one in flight), or -1
int nextCandidate(TuneClass& c, int inFlight) {
  std::lock_guard<std::mutex> lock(tuner().m);
  for (int i = 0; i < (int)c.score.size(); ++i)
    if (c.score[i] < 0 && i != inFlight) return i;
  return -1;
}

// records a measurement; true once the class is decided (E.sched is then its schedule)
bool recordCandidate(Engine& E, TuneClass& c, int i, double score) {
  Tuner& T = tuner();
  std::lock_guard<std::mutex> lock(T.m);
  if (c.best >= 0) {
    E.sched = c.sched;
    return true;
  }
  if (c.score[i] < 0 || score < c.score[i]) c.score[i] = score;
  int best = 0;
  for (int k = 0; k < (int)c.score.size(); ++k) {
    if (c.score[k] < 0) return false;
    if (c.score[k] < c.score[best]) best = k;
  }
  const bool single = E.nSlots == 1;
  c.best = best;
  c.sched = (single ? kOneSlot : kManySlots)[best];
  E.sched = c.sched;
  persist(E.tuneKey.substr(0, E.tuneKey.find(':')), E.tuneKey, c.sched, c.score[best]);
  return true;
}

const Schedule& candidate(const Engine& E, int i) { return (E.nSlots == 1 ? kOneSlot : kManySlots)[i]; }

}  // namespace

void GraphCache::reset() {
  if (exec) {
    if (launchedOn) (void)hipStreamSynchronize(launchedOn);
    (void)hipGraphExecDestroy(exec);
  }
  exec = nullptr;
  launchedOn = nullptr;
  key.clear();
}
void GraphCache::instantiate(hipGraph_t graph, const char* what) {
  reset();
  hipGraphExec_t x = nullptr;
  const hipError_t e = hipGraphInstantiate(&x, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) throw HipError{e, what};
  exec = x;
}
void GraphCache::launch(hipStream_t st) {
  DV_HIP(hipGraphLaunch(exec, st));
  launchedOn = st;
}
bool GraphCache::matches(const void* k, size_t n) const {
  return exec && key.size() == n && std::memcmp(key.data(), k, n) == 0;
}
void GraphCache::setKey(const void* k, size_t n) { key.assign((const char*)k, (const char*)k + n); }

void TileSet::release() {
  items.release();
  bodies.release();
  heavy.release();
  heavyIdx.release();
  counts.release();
  slots.release();
  hItems.clear();
  hBodies.clear();
  hSlots.clear();
}

void Engine::release() {
  table.release();
  active.release();
  admit.release();
  retire.release();
  out.release();
  states.release();
  views.release();
  if (outHost) (void)hipHostFree(outHost);
  outHost = nullptr;
  for (hipEvent_t& e : polled) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
  for (auto& pair : tev)
    for (hipEvent_t& e : pair) {
      if (e) (void)hipEventDestroy(e);
      e = nullptr;
    }
  outHostCap = 0;
  graph.reset();
  burst.reset();
  hostTable.clear();
  slotScene.clear();
  nSlots = 0;
}

namespace {

// Per-scene buffers and table entry.  Work buffers are sized by capacity, so
// that a captured graph stays valid while counts drift.
template <typename TM>
void prepareScene(Scene& s, Args<TM>& a) {
  constexpr bool kMixed = sizeof(TM) < sizeof(double);
  const long RsCap = std::max<long>(3L * s.ncCap + s.nbil, 1);
  const long itemsCap = (long)s.ncCap + s.nbil;
  const long tilesCap = std::max<long>(1, (itemsCap + kTC - 1) / kTC);
  const long bodyTiles = std::max<long>(1, ((long)s.nb + kTB - 1) / kTB);
  // a body reduced by a workgroup has more than kHeavyDegree of the 2 nc incidences
  const long lqCap = bodyTiles + 2L * std::max(s.ncCap, 1) / kHeavyDegree + 2;
  s.partials.reserve((size_t)4 * kSlots * tilesCap + 2 * kSlots + 2 * lqCap);
  s.work.reserve((size_t)5 * RsCap + 16);  // ring[4], gHat: FP64 in both solvers
  // TM block: rn[4], r, then J and JB (per contact), tmp and (mixed) the
  // four ring slots of D^T gamma (per body) of 4 components 32-byte aligned,
  // then (mixed) M^-1
  const size_t nbc = (size_t)std::max(s.nb, 1);
  const size_t tmElems = (size_t)5 * RsCap + 8 + 4 * (2 * (size_t)std::max(s.ncCap, 1) + nbc + (kMixed ? 4 * nbc : 0)) +
                         (kMixed ? nbc + 1 : 0) + 16;
  TM* WM;
  if (kMixed) {
    s.frame32.reserve((size_t)6 * std::max(s.ncCap, 1));  // n, u
    s.work32.reserve(tmElems);
    WM = reinterpret_cast<TM*>(s.work32.p);
  } else {
    s.workM64.reserve(tmElems);
    WM = reinterpret_cast<TM*>(s.workM64.p);
  }
  std::memset(&a, 0, sizeof(a));
  const double cnt = (double)((long)s.nc + s.nbil);
  a.dims.nc = s.nc; a.dims.nbil = s.nbil; a.dims.nb = s.nb; a.dims.maxIt = s.maxIt;
  a.dims.mu = s.mu; a.dims.tol = s.tol; a.dims.gdiff = 1.0 / (cnt * cnt);
  a.dims.igdiff = 1.0 / a.dims.gdiff;
  a.dims.warm = s.warm ? 1 : 0; a.dims.geomMode = s.geomMode;
  const long items = (long)s.nc + s.nbil;
  a.nTiles = (int)((items + kTC - 1) / kTC);
  a.nBodyTiles = (int)((s.nb + kTB - 1) / kTB);
  a.active = 1;
  a.frame64 = s.frame.p;
  a.cap = s.ncCap;
  a.cA = s.cA.p; a.cB = s.cB.p; a.iStart = s.iStart.p; a.cStart = s.cStart.p; a.posB = s.posB.p;
  a.bilIncStart = s.bilIncStart.p; a.bilInc = s.bilInc.p; a.bilA = s.bilA.p; a.bilB = s.bilB.p;
  a.heavy = s.heavy.p; a.counters = s.counters.p;
  a.minv64 = s.minv.p;
  a.gammaIO = s.gamma.p;
  a.rFlat = s.r.p;
  double* W = s.work.p;
  for (int k = 0; k < 4; ++k) a.ring[k] = W + k * RsCap;
  a.ringStride = RsCap;
  a.gHat = W + 4 * RsCap;
  for (int k = 0; k < 4; ++k) a.rn[k] = WM + k * RsCap;
  a.r = WM + 4 * RsCap;
  TM* after = WM + 5 * RsCap;
  uintptr_t p = (reinterpret_cast<uintptr_t>(after) + 31) & ~uintptr_t(31);
  a.J = reinterpret_cast<V4<TM>*>(p);
  a.JB = a.J + std::max<long>(s.ncCap, 1);
  a.tmp = a.JB + std::max<long>(s.ncCap, 1);
  if (kMixed) {
    a.sr = a.tmp + nbc;
    a.frame = reinterpret_cast<TM*>(s.frame32.p);
    a.minv = reinterpret_cast<TM*>(a.sr + 4 * nbc);
  } else {
    a.frame = reinterpret_cast<TM*>(s.frame.p);
    a.minv = reinterpret_cast<TM*>(s.minv.p);
  }
  for (int k = 0; k < 2; ++k) {
    a.part1[k] = s.partials.p + (long)(2 * k) * kSlots * tilesCap;
    a.part3[k] = s.partials.p + (long)(2 * k + 1) * kSlots * tilesCap;
    a.p1sum[k] = s.partials.p + 4L * kSlots * tilesCap + k * kSlots;
    a.lq[k] = s.partials.p + 4L * kSlots * tilesCap + 2 * kSlots + k * lqCap;
  }
}

template <typename TM>
Args<TM>& slotArgs(Engine& E, int slot) {
  return reinterpret_cast<Args<TM>*>(E.hostTable.data())[slot];
}

// Tiles of the slots in set.hSlots, uploaded with the device slot list and
// the heavy-body table (built on the device from the slots' counters).
template <typename TM>
void buildSet(Engine& E, TileSet& set, hipStream_t st) {
  set.hItems.clear();
  set.hBodies.clear();
  long heavyCap = 1;
  for (int slot : set.hSlots) {
    const Args<TM>& a = slotArgs<TM>(E, slot);
    for (int t = 0; t < a.nTiles; ++t) set.hItems.push_back(Tile{slot, t * kTC});
    for (int b = 0; b < a.dims.nb; b += kTB) set.hBodies.push_back(Tile{slot, b});
    heavyCap += 2L * std::max<long>(a.cap, 1) / kHeavyDegree + 1;
  }
  set.items.reserve(std::max<size_t>(set.hItems.size(), 1));
  set.bodies.reserve(std::max<size_t>(set.hBodies.size(), 1));
  set.heavy.reserve((size_t)heavyCap);
  set.heavyIdx.reserve((size_t)heavyCap);
  set.counts.reserve(4);
  set.slots.reserve(std::max<size_t>(set.hSlots.size(), 1));
  set.countsHost[0] = (int)set.hItems.size();
  set.countsHost[1] = (int)set.hBodies.size();
  set.countsHost[2] = 0;
  if (!set.hItems.empty())
    DV_HIP(hipMemcpyAsync(set.items.p, set.hItems.data(), sizeof(Tile) * set.hItems.size(),
                          hipMemcpyHostToDevice, st));
  if (!set.hBodies.empty())
    DV_HIP(hipMemcpyAsync(set.bodies.p, set.hBodies.data(), sizeof(Tile) * set.hBodies.size(),
                          hipMemcpyHostToDevice, st));
  if (!set.hSlots.empty())
    DV_HIP(hipMemcpyAsync(set.slots.p, set.hSlots.data(), sizeof(int) * set.hSlots.size(),
                          hipMemcpyHostToDevice, st));
  DV_HIP(hipMemcpyAsync(set.counts.p, set.countsHost, 3 * sizeof(int), hipMemcpyHostToDevice, st));
  if (!set.hSlots.empty())
    hipLaunchKernelGGL(kHeavyTable<TM>, dim3(1), dim3(1024), 0, st,
                       reinterpret_cast<const Args<TM>*>(E.table.p), set.slots.p, (int)set.hSlots.size(),
                       set.heavy.p, set.heavyIdx.p, set.counts.p);
}

template <typename TM>
Tables tablesOf(Engine& E, TileSet& set) {
  Tables T;
  // the padding too: a Plan holding these tables is compared bytewise with
  // the one a cached graph was captured with
  std::memset(&T, 0, sizeof(T));
  T.items = set.items.p;
  T.bodies = set.bodies.p;
  T.heavy = set.heavy.p;
  T.heavyIdx = set.heavyIdx.p;
  T.counts = set.counts.p;
  T.slots = set.slots.p;
  T.out = reinterpret_cast<ApgdState*>(E.out.p);
  T.st = reinterpret_cast<ApgdState*>(E.states.p);
  T.pv = reinterpret_cast<PassView*>(E.views.p);
  T.single = E.nSlots == 1 ? 1 : 0;
  T.xcd = E.sched.xcd;
  return T;
}

// Grid of the contact kernels.  A single slot gets one workgroup per row tile
// its contact capacity can hold (the capacity, not the count, so that the
// captured graph survives until a capacity grows; P1 relies on this grid
// covering every tile).  Tiles are 512 rows: with 1024-row tiles, 257 tiles
// on 256 CUs made one CU carry twice the rows of the others, and P1 took 35%
// longer at 263k contacts than at 262k; 256-row tiles balance better still
// but cost more in partial sums.  Past 512 tiles a CU holds three P1
// workgroups at once (kP1MinWaves).
template <typename TM>
void grids(Engine& E, const TileSet& set, int& gc, int& gb) {
  if (E.nSlots == 1) {
    const Args<TM>& a = slotArgs<TM>(E, 0);
    const long itemsCap = a.cap + a.dims.nbil;
    const long tilesCap = std::max<long>(1, (itemsCap + kTC - 1) / kTC);
    gc = (int)tilesCap;
    gb = (int)std::max<long>(1, std::min<long>((a.dims.nb + kTB - 1) / kTB, 4L * E.numCU));
  } else {
    // several slots: wgc and wgb workgroups per CU over the tile tables,
    // each taking a few tiles, so that the table and view reads that start
    // every tile overlap across many workgroups (512 scenes: the solve
    // 330 ms per batch step with 32 per CU instead of 386 ms with 2 and 4)
    gc = E.sched.wgc * E.numCU;
    gb = E.sched.wgb * E.numCU;
  }
  (void)set;
}

template <typename TM>
Plan<TM> activePlan(Engine& E) {
  Plan<TM> P;
  std::memset(&P, 0, sizeof(P));
  P.A = reinterpret_cast<const Args<TM>*>(E.table.p);
  P.T = tablesOf<TM>(E, E.active);
  grids<TM>(E, E.active, P.gc, P.gb);
  P.nSlots = E.nSlots;
  P.rpt = E.sched.rpt;
  return P;
}

template <typename TM>
void rebuildActive(Engine& E, hipStream_t st) {
  E.active.hSlots.clear();
  for (int i = 0; i < E.nSlots; ++i)
    if (slotArgs<TM>(E, i).active) E.active.hSlots.push_back(i);
  buildSet<TM>(E, E.active, st);
}

template <typename TM>
void uploadTable(Engine& E, hipStream_t st) {
  DV_HIP(hipMemcpyAsync(E.table.p, E.hostTable.data(), E.hostTable.size(), hipMemcpyHostToDevice, st));
}

template <typename TM>
void startAdmitted(Engine& E, hipStream_t st) {
  if (E.admit.hSlots.empty()) return;
  uploadTable<TM>(E, st);
  buildSet<TM>(E, E.admit, st);
  const Args<TM>* A = reinterpret_cast<const Args<TM>*>(E.table.p);
  const Tables T = tablesOf<TM>(E, E.admit);
  int gc, gb;
  grids<TM>(E, E.admit, gc, gb);
  const int n = (int)E.admit.hSlots.size();
  // setup (APGD.cu:195-223)
  hipLaunchKernelGGL(kSetup1<TM>, dim3(gc), dim3(kTC), 0, st, A, T);
  hipLaunchKernelGGL((kBody<TM, 0>), dim3(gb), dim3(kTB), 0, st, A, T, 0);
  hipLaunchKernelGGL(kSetup3<TM>, dim3(gc), dim3(kTC), 0, st, A, T);
  hipLaunchKernelGGL((kBody<TM, 1>), dim3(gb), dim3(kTB), 0, st, A, T, 0);
  hipLaunchKernelGGL(kSetup5<TM>, dim3(gc), dim3(kTC), 0, st, A, T);
  hipLaunchKernelGGL(kSetupDecide<TM>, dim3(n), dim3(kTD), 0, st, A, T);
  E.admit.hSlots.clear();
  rebuildActive<TM>(E, st);
}

// the captured graph of a chunk of passes under the current plan
template <typename TM>
void prepareChunk(Engine& E, hipStream_t st) {
  const Plan<TM> P = activePlan<TM>(E);
  if (E.noGraph || E.graph.matches(&P, sizeof(P))) return;
  E.graph.reset();
  hipGraph_t graph;
  DV_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  enqueuePasses<TM>(P, st, kPassesPerChunk, nullptr, 0, nullptr);
  DV_HIP(hipStreamEndCapture(st, &graph));
  E.graph.instantiate(graph, "hipGraphInstantiate(apgd passes)");
  E.graph.setKey(&P, sizeof(P));
}

template <typename TM>
void runChunk(Engine& E, hipStream_t st, hipEvent_t* ev, int prof) {
  if (E.noGraph || ev) {
    enqueuePasses<TM>(activePlan<TM>(E), st, kPassesPerChunk, ev, prof, &E.burst);
  } else {
    prepareChunk<TM>(E, st);
    E.graph.launch(st);
  }
  E.launched += kPassesPerChunk;
}

// gamma := gammaHat and the objective for the slots, then their statistics
template <typename TM>
void retireSlots(Engine& E, const std::vector<int>& slots, hipStream_t st) {
  if (slots.empty()) return;
  E.retire.hSlots = slots;
  buildSet<TM>(E, E.retire, st);
  const Args<TM>* A = reinterpret_cast<const Args<TM>*>(E.table.p);
  const Tables T = tablesOf<TM>(E, E.retire);
  int gc, gb;
  grids<TM>(E, E.retire, gc, gb);
  const int fin = 1;  // a chunk's last pass (parity 1) decides into st[1]
  hipLaunchKernelGGL(kFin1<TM>, dim3(gc), dim3(kTC), 0, st, A, T, fin);
  hipLaunchKernelGGL((kBody<TM, 3>), dim3(gb), dim3(kTB), 0, st, A, T, 0);
  hipLaunchKernelGGL(kFin3<TM>, dim3(gc), dim3(kTC), 0, st, A, T);
  hipLaunchKernelGGL(kFinDecide<TM>, dim3((int)slots.size()), dim3(kTD), 0, st, A, T, fin);
  ApgdState* outH = reinterpret_cast<ApgdState*>(E.outHost);
  DV_HIP(hipMemcpyAsync(outH, E.out.p, sizeof(ApgdState) * E.nSlots, hipMemcpyDeviceToHost, st));
  DV_HIP(hipStreamSynchronize(st));
  for (int slot : slots) {
    Scene& s = *E.slotScene[slot];
    const ApgdState& S = outH[slot];
    s.stats.iterations = S.it;
    s.stats.backtracks = S.backtracks;
    s.stats.residual = S.best;
    s.stats.objective = S.objective;
    s.stats.passes = S.passes;
    s.stats.rows_per_thread = E.sched.rpt;
    s.stats.xcd_order = E.sched.xcd;
    s.stats.wg_per_cu_contacts = E.nSlots > 1 ? E.sched.wgc : 0;
    s.stats.wg_per_cu_bodies = E.sched.wgb;
    s.stats.schedule_tuning = E.tuning;
    if (s.verbose) printf("  Iterations: %d Residual: %g\n", S.it, S.best);
    slotArgs<TM>(E, slot).active = 0;
    E.slotScene[slot] = nullptr;
  }
  uploadTable<TM>(E, st);
  rebuildActive<TM>(E, st);
}

template <typename TM>
void openT(Engine& E, int nSlots) {
  E.hostTable.assign(sizeof(Args<TM>) * nSlots, 0);  // every slot inactive
  E.slotScene.assign(nSlots, nullptr);
  E.table.reserve(sizeof(Args<TM>) * nSlots);
}

}  // namespace

void engineOpen(Engine& E, int nSlots, bool mixed, int numCU, bool noGraph, int forcedRpt) {
  if (nSlots < 1) throw DvError{DV_ERR_ARG, "an engine needs at least one slot"};
  E.forcedRpt = (forcedRpt == 1 || forcedRpt == 2) ? forcedRpt : 0;
  E.sched = nSlots == 1 ? kOneSlot[0] : kManySlots[0];
  if (E.forcedRpt) E.sched.rpt = E.forcedRpt;
  E.tuning = 0;
  if (E.mixed != (mixed ? 1 : 0) || E.nSlots != nSlots) E.graph.reset();
  E.mixed = mixed ? 1 : 0;
  E.nSlots = nSlots;
  E.numCU = numCU;
  E.noGraph = noGraph;
  E.launched = 0;
  E.admit.hSlots.clear();
  if (mixed) openT<float>(E, nSlots);
  else openT<double>(E, nSlots);
  E.out.reserve(sizeof(ApgdState) * nSlots);
  E.states.reserve(2 * sizeof(ApgdState) * nSlots);
  E.views.reserve(2 * sizeof(PassView) * nSlots);
  if (E.outHostCap < nSlots) {
    if (E.outHost) (void)hipHostFree(E.outHost);
    E.outHost = nullptr;
    DV_HIP(hipHostMalloc(&E.outHost, 2 * sizeof(ApgdState) * nSlots, hipHostMallocDefault));
    E.outHostCap = nSlots;
  }
}

void engineAdmit(Engine& E, int slot, Scene& s) {
  if (slot < 0 || slot >= E.nSlots || E.slotScene[slot]) throw DvError{DV_ERR_ARG, "engine slot not free"};
  if ((long)s.nc + s.nbil == 0) throw DvError{DV_ERR_ARG, "a scene without rows needs no solve"};
  if (E.mixed) prepareScene<float>(s, slotArgs<float>(E, slot));
  else prepareScene<double>(s, slotArgs<double>(E, slot));
  E.slotScene[slot] = &s;
  E.admit.hSlots.push_back(slot);
}

void engineStartAdmitted(Engine& E, hipStream_t st) {
  if (E.mixed) startAdmitted<float>(E, st);
  else startAdmitted<double>(E, st);
}

void engineRunChunk(Engine& E, hipStream_t st) {
  if (E.mixed) runChunk<float>(E, st, nullptr, 0);
  else runChunk<double>(E, st, nullptr, 0);
}

std::vector<int> enginePoll(Engine& E, hipStream_t st) {
  ApgdState* outH = reinterpret_cast<ApgdState*>(E.outHost);
  DV_HIP(hipMemcpyAsync(outH, E.out.p, sizeof(ApgdState) * E.nSlots, hipMemcpyDeviceToHost, st));
  DV_HIP(hipStreamSynchronize(st));
  std::vector<int> done;
  for (int i = 0; i < E.nSlots; ++i)
    if (E.slotScene[i] && outH[i].done) done.push_back(i);
  return done;
}

void engineRetire(Engine& E, const std::vector<int>& slots, hipStream_t st) {
  if (E.mixed) retireSlots<float>(E, slots, st);
  else retireSlots<double>(E, slots, st);
}

bool engineBusy(const Engine& E) {
  for (Scene* s : E.slotScene)
    if (s) return true;
  return false;
}

namespace {

// Admit all, run chunks until every scene is done, retire all.  With
// profiling (single scene), one chunk per solve is launched on the stream
// with events around the chosen kernel: the first chunk after the first
// whose entry state (the last state the host read back) has the momentum
// term on, so that the timed contact pass moves all of its operands.
template <typename TM>
void solveAll(Engine& E, const std::vector<Scene*>& scenes, hipStream_t st, int prof,
              std::vector<hipEvent_t>* profEv, dv_stats* profStats) {
  const int n = (int)scenes.size();
  for (int i = 0; i < n; ++i) engineAdmit(E, i, *scenes[i]);
  // the schedule: fixed, decided for the solve's class, or measured on this
  // solve's chunks (not while a kernel is being profiled)
  long rows = 0;
  std::vector<long> slotRows(n);
  for (int i = 0; i < n; ++i) rows += slotRows[i] = (long)scenes[i]->nc + scenes[i]->nbil;
  TuneClass* tc = chooseSchedule(E, rows, prof == 0);
  E.tuning = tc ? 1 : 0;
  if (tc)
    for (auto& pair : E.tev)
      for (hipEvent_t& e : pair)
        if (!e) DV_HIP(hipEventCreate(&e));
  E.chunkCand[0] = E.chunkCand[1] = -1;
  startAdmitted<TM>(E, st);
  hipEvent_t* ev = nullptr;
  if (prof && profEv) {
    while (profEv->size() < 2) {
      hipEvent_t e;
      DV_HIP(hipEventCreate(&e));
      profEv->push_back(e);
    }
    ev = profEv->data();
  }
  int maxIt = 1;
  for (Scene* s : scenes) maxIt = std::max(maxIt, s->maxIt);
  const long passCap = (long)maxIt * 64 + 4096;
  int passesBefore = 0;
  ApgdState* outH = reinterpret_cast<ApgdState*>(E.outHost);
  ApgdState* buf[2] = {outH, outH + E.nSlots};
  for (int b = 0; b < 2; ++b)
    if (!E.polled[b]) DV_HIP(hipEventCreateWithFlags(&E.polled[b], hipEventDisableTiming));
  // a chunk, then the copy of the states after it into buf[b]; while the
  // class is being measured, the chunk runs the next unmeasured candidate
  // schedule between two timing events
  auto launch = [&](int b, bool timed) {
    const int ci = tc ? nextCandidate(*tc, E.chunkCand[b ^ 1]) : -1;
    if (ci >= 0) {
      E.sched = candidate(E, ci);
      if (!timed) prepareChunk<TM>(E, st);  // capture outside the timed interval
      DV_HIP(hipEventRecord(E.tev[b][0], st));
    }
    E.chunkCand[b] = ci;
    runChunk<TM>(E, st, timed ? ev : nullptr, prof);
    if (ci >= 0) DV_HIP(hipEventRecord(E.tev[b][1], st));
    DV_HIP(hipMemcpyAsync(buf[b], E.out.p, sizeof(ApgdState) * E.nSlots, hipMemcpyDeviceToHost, st));
    DV_HIP(hipEventRecord(E.polled[b], st));
  };
  bool sampled = false, curTimed = false;
  int cur = 0;
  // a measured chunk counts when no slot finished during it: then every
  // active slot ran all of its passes
  int doneBefore = 0;
  // slots already retired while others iterate (several slots only)
  std::vector<char> retired(n, 0);
  int nRetired = 0;
  double activeRows = (double)rows;
  launch(cur, false);
  for (;;) {
    // The next chunk is queued before this one's states are read, so the GPU
    // does not wait for the host between chunks (after the solve ends, the
    // queued chunk's passes exit at once).  While the profiling sample is
    // still to be taken, the next chunk's entry state must be known first.
    const bool ahead = !(ev && !sampled);
    if (ahead) launch(cur ^ 1, false);
    DV_HIP(hipEventSynchronize(E.polled[cur]));
    const ApgdState* last = buf[cur];
    if (curTimed && profStats && last[0].passes > passesBefore) {  // the first pass did work
      float ms = 0;
      DV_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
      profStats->prof_kernel_ms += ms;
      profStats->prof_kernel_count += prof == 4 ? 1 : kProfBurst;
    }
    passesBefore = last[0].passes;
    int doneNow = 0;
    double rowsLeft = 0;
    for (int i = 0; i < n; ++i) {
      if (last[i].done) ++doneNow;
      else rowsLeft += (double)slotRows[i];
    }
    if (tc && E.chunkCand[cur] >= 0 && doneNow == doneBefore && doneNow < n) {
      float ms = 0;
      DV_HIP(hipEventElapsedTime(&ms, E.tev[cur][0], E.tev[cur][1]));
      const double score = 1e3 * ms / kPassesPerChunk / activeRows * 1e6;
      if (recordCandidate(E, *tc, E.chunkCand[cur], score)) tc = nullptr;
    }
    doneBefore = doneNow;
    activeRows = rowsLeft;
    if (doneNow == n) break;
    // Once a quarter of the slots have finished since the last retirement,
    // retire them while the others iterate: the pass kernels then loop over
    // the active slots' tiles only, instead of reading the view of every
    // finished slot's tiles on every pass.  A finished slot's state is
    // frozen, so nothing changes in any result.  The retirement waits for
    // the chunk queued ahead (it runs on the same stream) and rebuilds the
    // tile tables in place, so the captured chunk stays valid.
    if (n > 1 && doneNow - nRetired >= std::max(4, n / 4)) {
      std::vector<int> fin;
      for (int i = 0; i < n; ++i)
        if (last[i].done && !retired[i]) {
          fin.push_back(i);
          retired[i] = 1;
        }
      nRetired += (int)fin.size();
      retireSlots<TM>(E, fin, st);
    }
    if (E.launched > passCap) throw DvError{DV_ERR_DIVERGED, "APGD exceeded its pass budget"};
    cur ^= 1;
    curTimed = false;
    if (!ahead) {
      const bool timed = ev && !sampled && !last[0].restart && !last[0].done;
      sampled = sampled || timed;
      launch(cur, timed);
      curTimed = timed;
    }
  }
  std::vector<int> slots;
  for (int i = 0; i < n; ++i)
    if (!retired[i]) slots.push_back(i);
  retireSlots<TM>(E, slots, st);
}

}  // namespace

void solveFusedMany(Engine& E, const std::vector<Scene*>& scenes, bool mixed, hipStream_t st, int numCU,
                    bool noGraph) {
  if (scenes.empty()) return;
  engineOpen(E, (int)scenes.size(), mixed, numCU, noGraph, scenes[0]->rowsPerThread);
  if (mixed) solveAll<float>(E, scenes, st, 0, nullptr, nullptr);
  else solveAll<double>(E, scenes, st, 0, nullptr, nullptr);
}

namespace {

// N x by the kernels a solver pass runs, on the assembled scene alone: the
// setup takes gamma_0 = x (warm start from the staging copy of x) and stores
// J = F^T x with the store P1 uses; then P2 (kBody<TM, 2>) forms tmp = M^-1
// (sum_A J - sum_B J + bilateral rows of x) and kRowsN the rows
// (N x)_c = F_c (tmp_A - tmp_B) and the bilateral rows, as the next P1 forms
// them, into N of ring slot 1 (the pass's gamma_new slot, which gets a copy
// of x).  y receives them in the reference's flat layout, as TM values.
template <typename TM>
void passProductT(Scene& s, const double* xDev, double* yHost) {
  Engine& E = s.engine;
  const hipStream_t st = s.stream;
  engineOpen(E, 1, sizeof(TM) < sizeof(double), s.numCU, true, s.rowsPerThread == 2 ? 2 : 1);
  engineAdmit(E, 0, s);
  Args<TM>& a = slotArgs<TM>(E, 0);
  a.gammaIO = const_cast<double*>(xDev);  // read only: warm start from x
  a.dims.warm = 1;
  a.dims.maxIt = 1;  // a live state, so that P2 gathers
  startAdmitted<TM>(E, st);
  DV_HIP(hipMemcpyAsync(a.ring[1], a.ring[0], sizeof(double) * a.ringStride, hipMemcpyDeviceToDevice, st));
  const Plan<TM> P = activePlan<TM>(E);
  launchPass<TM>(P, st, 2, 0, false);  // kBody<TM, 2> on the view the setup left (nxt = 1)
  hipLaunchKernelGGL(kRowsN<TM>, dim3(P.gc), dim3(kTC), 0, st, P.A, P.T, 0);
  std::vector<TM> rn((size_t)a.ringStride);
  DV_HIP(hipMemcpyAsync(rn.data(), a.rn[1], sizeof(TM) * rn.size(), hipMemcpyDeviceToHost, st));
  DV_HIP(hipStreamSynchronize(st));
  DV_HIP(hipGetLastError());
  for (int c = 0; c < s.nc; ++c)
    for (int rr = 0; rr < 3; ++rr) yHost[s.nbil + 3L * c + rr] = (double)rn[(size_t)rr * a.cap + c];
  for (int q = 0; q < s.nbil; ++q) yHost[q] = (double)rn[3 * (size_t)a.cap + q];
  a.active = 0;
  E.slotScene[0] = nullptr;
}

}  // namespace

void passProduct(Scene& s, bool mixed, const double* xDev, double* yHost) {
  if ((long)s.nc + s.nbil == 0) return;
  if (mixed) passProductT<float>(s, xDev, yHost);
  else passProductT<double>(s, xDev, yHost);
}

void solveFused(Scene& s, bool mixed) {
  std::vector<Scene*> one{&s};
  engineOpen(s.engine, 1, mixed, s.numCU, s.noGraph, s.rowsPerThread);
  if (mixed) solveAll<float>(s.engine, one, s.stream, s.profKernel, &s.profEv, &s.stats);
  else solveAll<double>(s.engine, one, s.stream, s.profKernel, &s.profEv, &s.stats);
}

}  // namespace dv
