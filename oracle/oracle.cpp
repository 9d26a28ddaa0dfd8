// ORACLE -- test infrastructure only.
//
// A CPU, FP64 restatement of DVIDE's contact time step, used as the checker
// for the HIP implementation in dvide_amd/.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it; the product never does.
//
// It follows the reference's algorithm operation by operation (citations are
// to files of melanz/DVIDE):
//   System::initializeSystem / processConstraints   System.cu:303-340, 342-424
//   CollisionDetector (AABB, uniform-grid broadphase,
//     sphere-sphere / sphere-box narrowphase)      CollisionDetector.cu:10-185, 187-501, 547-665
//   buildAppliedImpulseVector / multiplyByMass     System.cu:1411-1417, 1502-1514
//   buildContactJacobian (+ bilateral rows)        System.cu:622-638, 676-767, 904-967, 1321-1409
//   buildSchurVector / buildStabilization          System.cu:1516-1544, 1561-1580
//   APGD::solve, project, getResidual              APGD.cu:47-122, 164-361
//   PGJ / PGS (secondary cross-check solvers)      PGJ.cu:66-81, 151-210; PGS.cu:149-252, 384-447
//   DoTimeStep update                              System.cu:544-602
//   getCCPViolation                                System.cu:1590-1643
//   ANCF shell meshes: importMesh, contact points,
//     their positions and 3x36 Jacobian blocks     System.cu:474-500, 676-767, 904-967, 1932-2060;
//                                                  CollisionDetector.cu:50-69, 226-280
//
// Linear algebra mirrors CUSP semantics: D and D^T are COO matrices sorted by
// row (D^T by a stable sort, as cusp::coo_matrix::sort_by_row), and a COO
// SpMV y = A x accumulates each row's entries in storage order.  Dot products
// and norms are summed in the fixed CR256 order (see cr256), so results do not
// depend on the OpenMP thread count.  Built with -ffp-contract=off so that no
// multiply-add is fused: the narrowphase keep test and contact frames are then
// evaluated exactly as the HIP kernels evaluate them.
//
// Pinned by (tests/test_oracle.py): brute-force all-pairs contact sets and
// closed-form sphere/box geometry, the closed form of the Coulomb-cone
// projection and a numerical nearest-point search, dense D M^-1 D^T products,
// free fall by semi-implicit Euler, the ball-on-box normal impulse m g h and
// the stick/slip threshold |g_x| = mu g, the pile weight balance, the CCP
// optimality conditions at convergence, and the velocities of converged
// projected Gauss-Seidel and Jacobi runs (another route to the same CCP).
// The shell element's shape functions are pinned by their corner values and
// slopes, the exact interpolation of a flat element, and the reference's
// expanded polynomials; the shell Jacobian blocks by rigid translations and
// by the dense product D q against the contact points' positions.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct V3 {
  double x, y, z;
};
inline V3 mk(double x, double y, double z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return mk(-a.x, -a.y, -a.z); }
inline V3 operator*(V3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(double s, V3 a) { return mk(s * a.x, s * a.y, s * a.z); }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// length / normalize.  The reference's double3 helpers call sqrtf/rsqrtf
// (helper_math_double.h:576-579, 594-598), which nvcc resolves to the binary32
// functions: lengths and unit normals carry binary32 rounding.  Mode 1
// restates that (correctly rounded binary32 sqrt and reciprocal, computed via
// double operations whose second rounding is innocuous); mode 0 is plain FP64.
int g_geomMode = 1;
inline double lengthOf(V3 v) {
  double d = dot(v, v);
  if (g_geomMode == 1) return (double)(float)std::sqrt((double)(float)d);
  return std::sqrt(d);
}
inline double invLengthOf(V3 v) {
  double d = dot(v, v);
  if (g_geomMode == 1) {
    float s = (float)std::sqrt((double)(float)d);
    return (double)(float)(1.0 / (double)s);
  }
  return 1.0 / std::sqrt(d);
}
inline V3 normalize(V3 v) { return v * invLengthOf(v); }

// Deterministic reductions in a fixed order ("CR256"): the vector is cut
// into consecutive blocks of 256 terms (zero-padded); each block is reduced
// by a binary tree with strides 128, 64, ..., 1 (t[i] += t[i + s] for i < s);
// the block results are then added left to right.  A 256-thread GPU block
// reproduces this order exactly, which makes the plain FP64 HIP path's dot
// products bit-identical to these.
template <class F>
double cr256(size_t n, F term) {
  const size_t nblk = (n + 255) / 256;
  std::vector<double> part(nblk, 0.0);
#pragma omp parallel for schedule(static)
  for (long blk = 0; blk < (long)nblk; ++blk) {
    double t[256];
    for (int i = 0; i < 256; ++i) {
      size_t q = (size_t)blk * 256 + i;
      t[i] = q < n ? term(q) : 0.0;
    }
    for (int st = 128; st >= 1; st >>= 1)
      for (int i = 0; i < st; ++i) t[i] = t[i] + t[i + st];
    part[blk] = t[0];
  }
  double s = 0;
  for (double p : part) s += p;
  return s;
}

double vdot(const std::vector<double>& a, const std::vector<double>& b) {
  return cr256(a.size(), [&](size_t i) { return a[i] * b[i]; });
}
double vnrm2(const std::vector<double>& a) {
  return std::sqrt(cr256(a.size(), [&](size_t i) { return a[i] * a[i]; }));
}
double vnrmmax(const std::vector<double>& a) {
  double m = 0;
  for (double x : a) m = std::max(m, std::fabs(x));
  return m;
}
// z = a*x + b*y   (cusp::blas::axpby)
void axpby(const std::vector<double>& x, const std::vector<double>& y, std::vector<double>& z,
           double a, double b) {
  z.resize(x.size());
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)x.size(); ++i) z[i] = a * x[i] + b * y[i];
}

// COO matrix sorted by row, with a row-pointer index for the SpMV.
struct Coo {
  int rows = 0, cols = 0;
  std::vector<int> I, J;
  std::vector<double> V;
  std::vector<size_t> rowPtr;
  void index() {
    rowPtr.assign(rows + 1, 0);
    for (int r : I) rowPtr[r + 1]++;
    for (int r = 0; r < rows; ++r) rowPtr[r + 1] += rowPtr[r];
  }
  // y = A x
  void mul(const std::vector<double>& x, std::vector<double>& y) const {
    y.assign(rows, 0.0);
#pragma omp parallel for schedule(static)
    for (long r = 0; r < rows; ++r) {
      double s = 0;
      for (size_t k = rowPtr[r]; k < rowPtr[r + 1]; ++k) s += V[k] * x[J[k]];
      y[r] = s;
    }
  }
};

struct Aabb {
  V3 lo, hi;
};

// Shape functions of the 4-node ANCF shell element (the incomplete cubic
// interpolation of System.cu:904-961 and CollisionDetector.cu:50-69) at
// (xi, eta) in [0,1]^2, for an element of side lengths a (along xi) and b
// (along eta).  Node q sits at corner (0,0), (1,0), (1,1), (0,1) for q = 0..3;
// S[3q] weights its position, S[3q+1] its x-gradient and S[3q+2] its
// y-gradient.  In the corner's own coordinates u = |xi - xi_q|, w = |eta -
// eta_q|: position (1-u)(1-w)(1+u+w-2u^2-2w^2), x-gradient +-a u(1-u)^2 (1-w),
// y-gradient +-b w(1-w)^2 (1-u), the sign following the direction of u, w.
void shellShape(double xi, double eta, double a, double b, double S[12]) {
  for (int q = 0; q < 4; ++q) {
    const bool fx = q == 1 || q == 2, fy = q >= 2;
    const double u = fx ? 1.0 - xi : xi, w = fy ? 1.0 - eta : eta;
    const double ou = 1.0 - u, ow = 1.0 - w;
    S[3 * q] = (ou * ow) * (((1.0 + u) + w) - 2.0 * (u * u) - 2.0 * (w * w));
    S[3 * q + 1] = (fx ? -a : a) * ((u * (ou * ou)) * ow);
    S[3 * q + 2] = (fy ? -b : b) * ((w * (ow * ow)) * ou);
  }
}

struct System {
  // parameters (System.cu:14-28, CollisionDetector.cu:503-514, APGD.cu:6-16)
  double h = 1e-3, mu = 0.25, tol = 1e-4, envelope = 1e-3;
  int maxIt = 100000;
  V3 gravity = mk(0, -9.81, 0);
  unsigned bins[3] = {20, 20, 20};
  bool warm = false, antiRelax = false;
  // Precision study of APGD (not a reference option): 0 FP64 throughout; 1
  // the Schur products N y, N gamma_new rounded to binary32 (the mixed GPU
  // solver's products, its iterates FP64); 2 also the iterates gamma_new and
  // y rounded to binary32.  Modes 1 and 2 take the Lipschitz test in the
  // cancellation-free form D' N D >= L |D|^2 (D = gamma_new - y), as the
  // mixed GPU solver does.
  int storage = 0;
  double time = 0;
  int timeIndex = 0;

  // bodies
  int nb = 0;
  std::vector<double> p, v, f, fEl, k, fContact, tmp, minv;  // minv per body DOF
  std::vector<double> geom;                                  // 3 per body
  std::vector<int> fixed, family;
  bool initialized = false;

  // ANCF shell mesh (System::importMesh, System.cu:1932-2060).  Its DOFs
  // follow the bodies': 9 per node (position, x-gradient, y-gradient).
  int nNodes = 0;
  std::vector<double> shellNodes;  // 9 per node
  std::vector<int> shellConn;      // 4 node indices per element
  std::vector<double> shellGeom;   // 4 per element: a, b, thickness, points per side
  std::vector<double> shellFext;   // 9 per node
  Coo shellMass;                   // M of the shell DOFs (rows and columns from 0)
  std::vector<int> shellInvI, shellInvJ;  // M^-1 triplets of the shell DOFs
  std::vector<double> shellInvV;
  // collision geometries: one per body, then each element's contact points
  // (System.cu:494-499: i, k = 1 .. points-2, xi = i/(points-1), eta = k/(points-1))
  int ng = 0;
  std::vector<int> ptEl, ptI, ptK;
  int ndof() const { return 3 * nb + 9 * nNodes; }

  // bilateral DOF constraints (System.cu:510-532)
  std::vector<int> bilA, bilB, bilOffset;
  std::vector<double> bilVel, bilT0, bilP0;
  int bilNnz = 0;

  // collision detection output
  std::vector<uint32_t> idA, idB;
  std::vector<double> nAndP;  // 4 per contact: -normal (B to A), -penetration
  std::vector<double> loc;    // 3 per contact
  size_t numPossible = 0;
  int nc = 0;

  // system of the step
  Coo D, DT, M;
  std::vector<double> friction, r, b, gamma;

  // solver outputs
  int iterations = 0;
  double residual = 0, objective = 0;
  int backtracks = 0;

  int rows() const { return 3 * nc + (int)bilA.size(); }

  // --- System::initializeSystem (bodies) and processConstraints -------------
  void initialize() {
    nb = (int)fixed.size();
    const int n3 = 3 * nb, nd = ndof();
    // the shell DOFs follow the bodies' in p, v, f (System.cu:440-466)
    p.resize(n3);
    v.resize(n3);
    p.insert(p.end(), shellNodes.begin(), shellNodes.end());
    v.resize(nd, 0.0);
    f.assign(nd, 0.0);
    for (int i = 0; i < 9 * nNodes; ++i) f[n3 + i] = shellFext[i];
    fEl.assign(nd, 0.0);  // the shells' elastic forces are not restated: zero
    k.assign(nd, 0.0);
    fContact.assign(nd, 0.0);
    tmp.assign(nd, 0.0);
    minv.resize(n3);
    // M^-1: the bodies' diagonal, then the shell triplets at offset 3 nb
    // (System.cu:466-472), sorted by row
    std::vector<int> I, J;
    std::vector<double> V;
    for (int j = 0; j < nb; ++j) {
      for (int c = 0; c < 3; ++c) {
        double g = c == 0 ? gravity.x : (c == 1 ? gravity.y : gravity.z);
        f[3 * j + c] = fixed[j] ? 0.0 : bodyMass[j] * g;
        I.push_back(3 * j + c);
        J.push_back(3 * j + c);
        V.push_back(fixed[j] ? 0.0 : 1.0 / bodyMass[j]);
        minv[3 * j + c] = V.back();
      }
    }
    for (size_t q = 0; q < shellInvI.size(); ++q) {
      I.push_back(n3 + shellInvI[q]);
      J.push_back(n3 + shellInvJ[q]);
      V.push_back(shellInvV[q]);
    }
    std::vector<size_t> perm(I.size());
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](size_t a, size_t b) { return I[a] < I[b]; });
    M = Coo();
    M.rows = M.cols = nd;
    for (size_t q : perm) {
      M.I.push_back(I[q]);
      M.J.push_back(J[q]);
      M.V.push_back(V[q]);
    }
    M.index();
    shellMass.rows = shellMass.cols = 9 * nNodes;
    shellMass.index();
    // collision geometries: the bodies', then the shells' contact points
    ptEl.clear(); ptI.clear(); ptK.clear();
    for (size_t e = 0; e < shellConn.size() / 4; ++e) {
      const int w = (int)shellGeom[4 * e + 3];
      for (int i = 1; i < w - 1; ++i)
        for (int kk = 1; kk < w - 1; ++kk) {
          ptEl.push_back((int)e);
          ptI.push_back(i);
          ptK.push_back(kk);
        }
    }
    ng = nb + (int)ptEl.size();
    gamma.assign(nd, 0.0);  // gamma_d = a_h (System.cu:185): zeros, one per DOF
    bilOffset.clear();
    bilP0.assign(bilA.size(), 0.0);
    int off = 0;
    for (size_t i = 0; i < bilA.size(); ++i) {
      bilOffset.push_back(off);
      if (bilB[i] < 0) {
        off += 1;
        bilP0[i] = p[bilA[i]];
      } else {
        off += 2;
      }
    }
    bilNnz = off;
    initialized = true;
  }
  std::vector<double> bodyMass;

  V3 bodyPos(int g) const { return mk(p[3 * g], p[3 * g + 1], p[3 * g + 2]); }

  // contact point j of the shells: its element's shape functions at (xi, eta)
  void pointShape(int j, double S[12]) const {
    const int e = ptEl[j];
    const double w1 = shellGeom[4 * e + 3] - 1.0;
    shellShape((double)ptI[j] / w1, (double)ptK[j] / w1, shellGeom[4 * e], shellGeom[4 * e + 1], S);
  }
  // its position from the element's nodal coordinates in p
  // (CollisionDetector.cu:50-69), summed node by node
  V3 pointPos(int j) const {
    double S[12];
    pointShape(j, S);
    const int e = ptEl[j];
    double x[3] = {0, 0, 0};
    for (int q = 0; q < 4; ++q) {
      const double* nd = &p[3 * (size_t)nb + 9 * (size_t)shellConn[4 * e + q]];
      for (int c = 0; c < 3; ++c) x[c] += (S[3 * q] * nd[c] + S[3 * q + 1] * nd[3 + c]) + S[3 * q + 2] * nd[6 + c];
    }
    return mk(x[0], x[1], x[2]);
  }
  // collision geometry g: a body's, or a shell contact point (a sphere of
  // half the thickness, collision family -2, System.cu:494-499)
  V3 geomPos(int g) const { return g < nb ? bodyPos(g) : pointPos(g - nb); }
  V3 geomVec(int g) const {
    if (g < nb) return mk(geom[3 * g], geom[3 * g + 1], geom[3 * g + 2]);
    return mk(0.5 * shellGeom[4 * ptEl[g - nb] + 2], 0, 0);
  }
  int famOf(int g) const { return g < nb ? family[g] : -2; }
  int ownerOf(int g) const { return g < nb ? g : nb + ptEl[g - nb]; }  // collisionMap.x
  V3 halfExt(int g) const {
    V3 e = geomVec(g);
    if (e.y == 0) e = mk(e.x, e.x, e.x);
    return e;
  }

  // --- generateAabbData (CollisionDetector.cu:24-79) ------------------------
  std::vector<Aabb> aabbs() const {
    std::vector<Aabb> a(ng);
    for (int g = 0; g < ng; ++g) {
      V3 pos = geomPos(g), e = halfExt(g);
      e = e + mk(envelope, envelope, envelope);
      a[g].lo = pos - e;
      a[g].hi = pos + e;
    }
    return a;
  }

  // --- detectPossibleCollisions_spatialSubdivision (CollisionDetector.cu:555-631)
  std::vector<uint64_t> broadphase() {
    std::vector<Aabb> a = aabbs();
    std::vector<uint64_t> keys;
    if (ng == 0) return keys;
    V3 lo = a[0].lo, hi = a[0].lo;
    for (const Aabb& b : a) {
      for (V3 q : {b.lo, b.hi}) {
        lo = mk(std::min(lo.x, q.x), std::min(lo.y, q.y), std::min(lo.z, q.z));
        hi = mk(std::max(hi.x, q.x), std::max(hi.y, q.y), std::max(hi.z, q.z));
      }
    }
    lo = lo - mk(0.01, 0.01, 0.01);
    hi = hi + mk(0.01, 0.01, 0.01);
    V3 bsi = mk((double)bins[0] / std::fabs(hi.x - lo.x), (double)bins[1] / std::fabs(hi.y - lo.y),
                (double)bins[2] / std::fabs(hi.z - lo.z));
    for (Aabb& b : a) {
      b.lo = b.lo - lo;
      b.hi = b.hi - lo;
    }
    // (bin, aabb) pairs, then group by bin
    std::vector<std::pair<uint32_t, uint32_t>> bi;
    for (int g = 0; g < ng; ++g) {
      unsigned x0 = (unsigned)(a[g].lo.x * bsi.x), y0 = (unsigned)(a[g].lo.y * bsi.y),
               z0 = (unsigned)(a[g].lo.z * bsi.z);
      unsigned x1 = (unsigned)(a[g].hi.x * bsi.x), y1 = (unsigned)(a[g].hi.y * bsi.y),
               z1 = (unsigned)(a[g].hi.z * bsi.z);
      for (unsigned i = x0; i <= x1; ++i)
        for (unsigned j = y0; j <= y1; ++j)
          for (unsigned kk = z0; kk <= z1; ++kk)
            bi.push_back({i + j * bins[0] + kk * bins[0] * bins[1], (uint32_t)g});
    }
    std::stable_sort(bi.begin(), bi.end(),
                     [](const std::pair<uint32_t, uint32_t>& x,
                        const std::pair<uint32_t, uint32_t>& y) { return x.first < y.first; });
    size_t s = 0;
    while (s < bi.size()) {
      size_t e = s;
      while (e < bi.size() && bi[e].first == bi[s].first) ++e;
      for (size_t i = s; i < e; ++i) {
        uint32_t ga = bi[i].second;
        const Aabb& A = a[ga];
        for (size_t kk = i + 1; kk < e; ++kk) {
          uint32_t gb = bi[kk].second;
          const Aabb& B = a[gb];
          bool in = (A.lo.x <= B.hi.x && B.lo.x <= A.hi.x) &&
                    (A.lo.y <= B.hi.y && B.lo.y <= A.hi.y) &&
                    (A.lo.z <= B.hi.z && B.lo.z <= A.hi.z);
          if (ownerOf(ga) == ownerOf(gb)) in = false;  // same item (collisionMap.x equal)
          if (famOf(ga) == famOf(gb) && famOf(ga) != -1) in = false;
          if (in) {
            uint64_t lo32 = std::min(ga, gb), hi32 = std::max(ga, gb);
            keys.push_back((lo32 << 32) | hi32);
          }
        }
      }
      s = e;
    }
    std::stable_sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    return keys;
  }

  // --- narrowphase keep test and contact record
  //     (countActualCollisions CollisionDetector.cu:284-328,
  //      storeActualCollisions CollisionDetector.cu:432-498)
  bool narrow(int ga, int gb, double* nOut, double* pen, V3* locOut) const {
    V3 posA = geomPos(ga), posB = geomPos(gb);
    V3 gA = geomVec(ga), gB = geomVec(gb);
    V3 normal = mk(1, 0, 0), location = mk(0, 0, 0);
    double penetration = 0;
    if (gA.y == 0 && gB.y == 0) {
      penetration = (gA.x + gB.x) - lengthOf(posB - posA);
      if (!(penetration >= -envelope)) return false;
      normal = normalize(posB - posA);  // from A to B
      location = posA + normal * gA.x;
    } else if ((gA.y != 0 && gB.y == 0) || (gA.y == 0 && gB.y != 0)) {
      double dmin = 0, rr = gB.x;
      V3 c = posB, bmin = posA - gA, bmax = posA + gA;
      if (gA.y == 0 && gB.y != 0) {
        rr = gA.x;
        c = posA;
        bmin = posB - gB;
        bmax = posB + gB;
      }
      V3 nn = mk(0, 0, 0);
      if (c.x < bmin.x) { dmin += (c.x - bmin.x) * (c.x - bmin.x); nn.x += c.x - bmin.x; }
      else if (c.x > bmax.x) { dmin += (c.x - bmax.x) * (c.x - bmax.x); nn.x += c.x - bmax.x; }
      if (c.y < bmin.y) { dmin += (c.y - bmin.y) * (c.y - bmin.y); nn.y += c.y - bmin.y; }
      else if (c.y > bmax.y) { dmin += (c.y - bmax.y) * (c.y - bmax.y); nn.y += c.y - bmax.y; }
      if (c.z < bmin.z) { dmin += (c.z - bmin.z) * (c.z - bmin.z); nn.z += c.z - bmin.z; }
      else if (c.z > bmax.z) { dmin += (c.z - bmax.z) * (c.z - bmax.z); nn.z += c.z - bmax.z; }
      penetration = rr - std::sqrt(dmin);
      if (!(penetration >= -envelope)) return false;
      normal = normalize(nn);
      location = c - rr * normal;
      if (gA.y == 0 && gB.y != 0) normal = -normal;
    } else {
      return false;  // box-box: no narrowphase
    }
    if (nOut) {
      nOut[0] = -normal.x; nOut[1] = -normal.y; nOut[2] = -normal.z;  // from B to A
      *pen = -penetration;                                             // the gap
      *locOut = location;
    }
    return true;
  }

  void collide() {
    std::vector<uint64_t> keys = broadphase();
    numPossible = keys.size();
    idA.clear(); idB.clear(); nAndP.clear(); loc.clear();
    std::vector<char> keep(keys.size());
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)keys.size(); ++i)
      keep[i] = narrow((int)(keys[i] >> 32), (int)(keys[i] & 0xffffffffu), nullptr, nullptr, nullptr);
    for (size_t i = 0; i < keys.size(); ++i) {
      if (!keep[i]) continue;
      int ga = (int)(keys[i] >> 32), gb = (int)(keys[i] & 0xffffffffu);
      double n[3], w;
      V3 l;
      narrow(ga, gb, n, &w, &l);
      idA.push_back(ga);
      idB.push_back(gb);
      nAndP.insert(nAndP.end(), {n[0], n[1], n[2], w});
      loc.insert(loc.end(), {l.x, l.y, l.z});
    }
    nc = (int)idA.size();
  }

  // --- buildAppliedImpulseVector (System.cu:1502-1514, 1411-1417) ------------
  void buildK() {
    const int n3 = 3 * nb, nd = ndof();
    for (int i = 0; i < n3; ++i) {
      double m = minv[i];
      if (m != 0) m = 1.0 / m;
      k[i] = m * v[i];
    }
    if (nNodes) {  // the shells' M v (System.cu:1510)
      std::vector<double> vs(v.begin() + n3, v.end()), ks;
      shellMass.mul(vs, ks);
      for (int i = 0; i < 9 * nNodes; ++i) k[n3 + i] = ks[i];
    }
    for (int i = 0; i < nd; ++i) k[i] = (h * f[i] + (-h) * fEl[i]) + 1.0 * k[i];
  }

  // --- contact frame (System.cu:753-762) -------------------------------------
  static void frame(const double* n3, V3* u, V3* w) {
    V3 n = mk(n3[0], n3[1], n3[2]);
    V3 uu;
    if (n.z != 0) uu = normalize(mk(1, 0, -n.x / n.z));
    else if (n.x != 0) uu = normalize(mk(-n.z / n.x, 0, 1));
    else uu = normalize(mk(1, -n.x / n.y, 0));
    *u = uu;
    *w = normalize(cross(n, uu));
  }

  // --- buildContactJacobian + transpose (System.cu:622-638, 676-1257, 1351-1409)
  // columns and weights of one side of a contact row: a body's 3 DOFs, or
  // the 36 DOFs of a shell point's element, node by node (position,
  // x-gradient, y-gradient, each x, y, z), weighted by the shape functions
  // (System.cu:785-800, 904-961)
  int sideEntries(int g, int* cols, double* wts) const {
    if (g < nb) {
      for (int q = 0; q < 3; ++q) { cols[q] = 3 * g + q; wts[q] = 1.0; }
      return 3;
    }
    double S[12];
    pointShape(g - nb, S);
    const int e = ptEl[g - nb];
    for (int q = 0; q < 4; ++q)
      for (int m = 0; m < 9; ++m) {
        cols[9 * q + m] = 3 * nb + 9 * shellConn[4 * e + q] + m;
        wts[9 * q + m] = S[3 * q + m / 3];
      }
    return 36;
  }
  int sideWidth(int g) const { return g < nb ? 3 : 36; }

  void buildD() {
    const int nbil = (int)bilA.size();
    D = Coo();
    D.rows = 3 * nc + nbil;
    D.cols = ndof();
    // each contact row holds its A side's entries, then its B side's
    std::vector<size_t> off(nc + 1, (size_t)bilNnz);
    for (int c = 0; c < nc; ++c) off[c + 1] = off[c] + 3 * (size_t)(sideWidth(idA[c]) + sideWidth(idB[c]));
    const size_t nnz = off[nc];
    D.I.resize(nnz); D.J.resize(nnz); D.V.resize(nnz);
    for (int i = 0; i < nbil; ++i) {
      int o = bilOffset[i];
      D.I[o] = i; D.J[o] = bilA[i]; D.V[o] = 1.0;
      if (bilB[i] >= 0) { D.I[o + 1] = i; D.J[o + 1] = bilB[i]; D.V[o + 1] = -1.0; }
    }
    friction.assign(nc, mu);
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nc; ++c) {
      size_t o = off[c];
      V3 u, w;
      frame(&nAndP[4 * c], &u, &w);
      V3 n = mk(nAndP[4 * c], nAndP[4 * c + 1], nAndP[4 * c + 2]);
      const V3 rowsv[3] = {n, u, w};
      int colA[36], colB[36];
      double wA[36], wB[36];
      const int na = sideEntries((int)idA[c], colA, wA), nbs = sideEntries((int)idB[c], colB, wB);
      for (int rr = 0; rr < 3; ++rr) {
        int row = 3 * (int)c + rr + nbil;
        const double vals[3] = {rowsv[rr].x, rowsv[rr].y, rowsv[rr].z};
        for (int q = 0; q < na; ++q) {
          D.I[o] = row; D.J[o] = colA[q]; D.V[o] = na == 3 ? vals[q] : wA[q] * vals[q % 3]; ++o;
        }
        for (int q = 0; q < nbs; ++q) {
          D.I[o] = row; D.J[o] = colB[q]; D.V[o] = nbs == 3 ? -vals[q] : -(wB[q] * vals[q % 3]); ++o;
        }
      }
    }
    D.index();
    // D^T: swap indices and stable-sort by row
    DT = Coo();
    DT.rows = D.cols;
    DT.cols = D.rows;
    std::vector<size_t> perm(nnz);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](size_t a, size_t b) { return D.J[a] < D.J[b]; });
    DT.I.resize(nnz); DT.J.resize(nnz); DT.V.resize(nnz);
    for (size_t q = 0; q < nnz; ++q) {
      DT.I[q] = D.J[perm[q]];
      DT.J[q] = D.I[perm[q]];
      DT.V[q] = D.V[perm[q]];
    }
    DT.index();
  }

  // --- buildSchurVector (System.cu:1516-1544, 1561-1580) ---------------------
  void buildR() {
    const int nbil = (int)bilA.size();
    M.mul(k, tmp);
    D.mul(tmp, r);
    b.assign(rows(), 0.0);
    for (int i = 0; i < nbil; ++i) {
      double vel = bilVel[i], t0 = bilT0[i];
      if (time < t0) vel = 0;
      double viol = bilB[i] < 0 ? p[bilA[i]] - bilP0[i] - vel * (time - t0)
                                : p[bilA[i]] - p[bilB[i]] - vel * (time - t0);
      b[i] = viol / h;
    }
    for (int c = 0; c < nc; ++c) {
      b[3 * c + nbil] = nAndP[4 * c + 3] / h;
      b[3 * c + 1 + nbil] = 0;
      b[3 * c + 2 + nbil] = 0;
    }
    for (size_t i = 0; i < r.size(); ++i) r[i] = 1.0 * b[i] + r[i];
  }

  // N x = D (M^-1 (D^T x))  (APGD.cu:81-87)
  void schur(const std::vector<double>& x, std::vector<double>& y) {
    DT.mul(x, fContact);
    M.mul(fContact, tmp);
    D.mul(tmp, y);
  }

  // project (APGD.cu:47-78) on a flat vector, contacts at offset `off`
  static void projectOne(double* g, double m) {
    double gn = g[0];
    double gt = std::sqrt(g[1] * g[1] + g[2] * g[2]);
    if (m == 0) {
      g[1] = 0; g[2] = 0;
      if (gn < 0) g[0] = 0;
    } else if (gt < m * gn) {
      // inside the cone
    } else if ((gt < -(1.0 / m) * gn) || (std::fabs(gn) < 10e-15)) {
      g[0] = 0; g[1] = 0; g[2] = 0;
    } else {
      double gnp = (gt * m + gn) / (m * m + 1.0);
      double gtp = gnp * m;
      double q = gtp / gt;
      g[0] = gnp; g[1] = q * g[1]; g[2] = q * g[2];
    }
  }
  void project(std::vector<double>& x) const {
    const int off = (int)bilA.size();
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nc; ++c) projectOne(&x[3 * c + off], friction[c]);
  }

  // getResidual (APGD.cu:90-102)
  double getResidual(const std::vector<double>& src, std::vector<double>& t) {
    double cnt = (double)((size_t)nc + bilA.size());
    double gdiff = 1.0 / (cnt * cnt);
    schur(src, t);
    for (size_t i = 0; i < t.size(); ++i) t[i] = 1.0 * r[i] + t[i];
    if (nc) {
      axpby(src, t, t, 1.0, -gdiff);
      project(t);
      axpby(src, t, t, 1.0 / gdiff, -1.0 / gdiff);
    }
    return vnrmmax(t);
  }

  // APGD::solve (APGD.cu:164-361)
  void apgd() {
    const size_t R = (size_t)rows();
    gamma.resize(R, 0.0);  // thrust resize: keeps the leading values, zero-fills
    std::vector<double> gammaHat(R, 1.0), gammaNew(R), g(R), y(R), yNew(R), t(R);
    if (!warm)
      for (int c = 0; c < nc; ++c) { gamma[3 * c] = 1.0; gamma[3 * c + 1] = 0; gamma[3 * c + 2] = 0; }
    y = gamma;
    double theta = 1.0, thetaNew = 1.0, beta = 0, obj1 = 0, obj2 = 0, best = 10e30;
    axpby(gamma, gammaHat, t, 1.0, -1.0);
    double L = vnrm2(t);
    std::vector<double> nt;
    schur(t, nt);
    L = vnrm2(nt) / L;
    double tt = 1.0 / L;
    backtracks = 0;
    int it;
    std::vector<double> Ny, Ng;
    // precision study (storage 1, 2): binary32 rounding of products / iterates
    auto round32 = [](std::vector<double>& x) {
      for (double& e : x) e = (double)(float)e;
    };
    auto product = [&](const std::vector<double>& x, std::vector<double>& out) {
      schur(x, out);
      if (storage >= 1) round32(out);
    };
    auto step = [&]() {
      axpby(y, g, gammaNew, 1.0, -tt);
      project(gammaNew);
      if (storage == 2) round32(gammaNew);
      product(gammaNew, Ng);
    };
    // the Lipschitz test (APGD.cu:236-264), true when the step is rejected
    auto rejected = [&]() {
      axpby(gammaNew, y, t, 1.0, -1.0);
      if (storage >= 1) return vdot(t, Ng) - vdot(t, Ny) >= L * vdot(t, t);
      obj1 = 0.5 * vdot(gammaNew, Ng) + vdot(gammaNew, r);
      obj2 = 0.5 * vdot(y, Ny) + vdot(y, r);
      const double nr = vnrm2(t);
      obj2 += vdot(g, t) + 0.5 * L * (nr * nr);
      return obj1 >= obj2;
    };
    for (it = 0; it < maxIt; ++it) {
      product(y, Ny);
      axpby(Ny, r, g, 1.0, 1.0);
      step();
      while (rejected()) {
        ++backtracks;
        L = 2.0 * L;
        tt = 1.0 / L;
        step();
      }
      thetaNew = (-(theta * theta) + theta * std::sqrt(theta * theta + 4.0)) / 2.0;
      beta = theta * (1.0 - theta) / (theta * theta + thetaNew);
      axpby(gammaNew, gamma, yNew, 1.0 + beta, -beta);
      if (storage == 2) round32(yNew);
      double res = getResidual(gammaNew, t);
      if (res < best) {
        best = res;
        gammaHat = gammaNew;
      }
      if (best < tol) break;
      axpby(gammaNew, gamma, t, 1.0, -1.0);
      if (vdot(g, t) > 0) {
        yNew = gammaNew;
        thetaNew = 1.0;
      }
      L = 0.9 * L;
      tt = 1.0 / L;
      theta = thetaNew;
      gamma = gammaNew;
      y = yNew;
      if (antiRelax) {
        // APGD.cu:335-345
        for (size_t i = 0; i < R; ++i) r[i] = -1.0 * antiRelaxVec[i] + r[i];
        DT.mul(gamma, fContact);
        for (int i = 0; i < ndof(); ++i) tmp[i] = k[i] + fContact[i];
        M.mul(tmp, v);
        D.mul(v, t);
        const int off = (int)bilA.size();
        (void)off;
        // updateAntiRelaxationVector indexes s and antiRelaxation without the
        // bilateral offset (APGD.cu:104-114)
        for (int c = 0; c < nc; ++c) {
          double sv = t[3 * c + 1], sw = t[3 * c + 2];
          antiRelaxVec[3 * c] = std::sqrt(sv * sv + sw * sw) * friction[c];
          antiRelaxVec[3 * c + 1] = 0;
          antiRelaxVec[3 * c + 2] = 0;
        }
        for (size_t i = 0; i < R; ++i) r[i] = 1.0 * antiRelaxVec[i] + r[i];
      }
    }
    schur(gammaHat, t);
    objective = 0.5 * vdot(gammaHat, t) + vdot(gammaHat, r);
    iterations = it;
    residual = best;
    gamma = gammaHat;
  }
  std::vector<double> antiRelaxVec;

  // --- projected Jacobi and Gauss-Seidel (PGJ.cu:151-210, PGS.cu:149-233,
  // 236-252, 384-447): the reference's secondary solvers of the same cone
  // complementarity problem.  They reach the CCP solution by another route,
  // so the velocities they give (unique, unlike gamma) cross-check APGD's.
  // Like the reference they handle contact rows only (no bilateral rows).
  int solverKind = 1;  // System(int solverType) numbering: 1 APGD, 5 PGJ, 6 PGS

  // per-contact step scale B_c = 3 / ((m_A^-1 + m_B^-1) |[n; u; v]|^2)
  // (PGS.cu:236-252, PGJ.cu:66-81), from the body-A block of D
  void pgScale(std::vector<double>& B) const {
    B.assign(3 * (size_t)nc, 0.0);
    for (int c = 0; c < nc; ++c) {
      const double* d = &D.V[18 * (size_t)c];
      const double mA = minv[3 * (size_t)idA[c]], mB = minv[3 * (size_t)idB[c]];
      double sq = 0;
      const int at[9] = {0, 1, 2, 6, 7, 8, 12, 13, 14};
      for (int q = 0; q < 9; ++q) sq += d[at[q]] * d[at[q]];
      const double g = 3.0 / ((mA + mB) * sq);
      B[3 * c] = B[3 * c + 1] = B[3 * c + 2] = g;
    }
  }

  // stopping measure of both (PGS.cu:423-441): the largest of the cone
  // infeasibility of gamma, the dual-cone infeasibility of w = N gamma + r
  // and the per-contact complementarity |gamma_c . w_c|
  double pgResidual(std::vector<double>& w) {
    schur(gamma, w);
    for (size_t i = 0; i < w.size(); ++i) w[i] = 1.0 * r[i] + w[i];
    double fx = 0, fy = 0, comp = 0;
    for (int c = 0; c < nc; ++c) {
      const double m = friction[c];
      const double* g = &gamma[3 * (size_t)c];
      const double* x = &w[3 * (size_t)c];
      double a = m * g[0] - std::sqrt(g[1] * g[1] + g[2] * g[2]);
      if (a != a) a = 0.0;
      fx = std::max(fx, -std::min(0.0, a));
      double bb = (1.0 / m) * x[0] - std::sqrt(x[1] * x[1] + x[2] * x[2]);
      if (bb != bb) bb = 0.0;
      fy = std::max(fy, -std::min(0.0, bb));
      comp = std::max(comp, std::fabs(x[0] * g[0] + x[1] * g[1] + x[2] * g[2]));
    }
    return std::max(std::max(fx, fy), comp);
  }

  void projectedIterations() {
    const double omega = 0.3, lambda = 2.0 / 3.0;  // PGS.cu:13-14, PGJ.cu:14-15
    iterations = -1;
    if (!bilA.empty() || nNodes) {  // contact rows between bodies only
      gamma.assign(rows(), 0.0);     // refused: the step goes on with no impulses
      return;
    }
    gamma.assign(3 * (size_t)nc, 0.0);
    for (int c = 0; c < nc; ++c) gamma[3 * c] = 1.0;
    std::vector<double> B, w, gh;
    pgScale(B);
    std::vector<double> vel((size_t)ndof());
    if (solverKind == 6) {  // v = M^-1 (D^T gamma) + M^-1 k
      DT.mul(gamma, fContact);
      M.mul(fContact, vel);
      M.mul(k, tmp);
      for (size_t i = 0; i < vel.size(); ++i) vel[i] = 1.0 * tmp[i] + vel[i];
    }
    double res = 0;
    int it;
    for (it = 0; it < maxIt; ++it) {
      if (solverKind == 6) {
        // one Gauss-Seidel sweep in contact order, velocities updated in place
        for (int c = 0; c < nc; ++c) {
          const double* d = &D.V[18 * (size_t)c];
          const V3 n = mk(d[0], d[1], d[2]), u = mk(d[6], d[7], d[8]), t = mk(d[12], d[13], d[14]);
          const size_t a = 3 * (size_t)idA[c], bq = 3 * (size_t)idB[c];
          const V3 vA = mk(vel[a], vel[a + 1], vel[a + 2]), vB = mk(vel[bq], vel[bq + 1], vel[bq + 2]);
          double* g = &gamma[3 * (size_t)c];
          const double old[3] = {g[0], g[1], g[2]};
          double x[3];
          x[0] = old[0] - omega * B[3 * c] * (dot(n, vA) - dot(n, vB) + b[3 * c]);
          x[1] = old[1] - omega * B[3 * c + 1] * (dot(u, vA) - dot(u, vB));
          x[2] = old[2] - omega * B[3 * c + 2] * (dot(t, vA) - dot(t, vB));
          projectOne(x, friction[c]);
          double dl[3];
          for (int q = 0; q < 3; ++q) {
            g[q] = lambda * x[q] + (1.0 - lambda) * old[q];
            dl[q] = lambda * (x[q] - old[q]);
          }
          const V3 imp = mk(n.x * dl[0] + u.x * dl[1] + t.x * dl[2], n.y * dl[0] + u.y * dl[1] + t.y * dl[2],
                            n.z * dl[0] + u.z * dl[1] + t.z * dl[2]);
          const double mA = minv[a], mB = minv[bq];
          vel[a] = vA.x + mA * imp.x; vel[a + 1] = vA.y + mA * imp.y; vel[a + 2] = vA.z + mA * imp.z;
          vel[bq] = vB.x - mB * imp.x; vel[bq + 1] = vB.y - mB * imp.y; vel[bq + 2] = vB.z - mB * imp.z;
        }
      } else {
        // one Jacobi step: gamma := lambda Pi(gamma - omega B (N gamma + r)) + (1 - lambda) gamma
        schur(gamma, w);
        gh.resize(gamma.size());
        for (size_t i = 0; i < gamma.size(); ++i) gh[i] = gamma[i] - omega * (B[i] * (w[i] + r[i]));
        project(gh);
        for (size_t i = 0; i < gamma.size(); ++i) gamma[i] = lambda * gh[i] + (1.0 - lambda) * gamma[i];
      }
      res = pgResidual(w);
      if (res < tol) break;
    }
    iterations = it;
    residual = res;
  }

  // --- System::DoTimeStep (System.cu:544-602) --------------------------------
  void step() {
    collide();
    buildK();
    const int nbil = (int)bilA.size();
    if (nc || nbil) {
      buildD();
      buildR();
      antiRelaxVec.assign(rows(), 0.0);
      if (solverKind == 5 || solverKind == 6) projectedIterations();
      else apgd();
      DT.mul(gamma, fContact);
      for (int i = 0; i < ndof(); ++i) tmp[i] = 1.0 * k[i] + 1.0 * fContact[i];
      M.mul(tmp, v);
      for (int i = 0; i < ndof(); ++i) fContact[i] = (1.0 / h) * fContact[i];
    } else {
      M.mul(k, v);
      std::fill(fContact.begin(), fContact.end(), 0.0);
      iterations = 0;
      residual = 0;
    }
    for (int i = 0; i < ndof(); ++i) p[i] = h * v[i] + p[i];
    time += h;
    timeIndex++;
  }

  // --- getCCPViolation (System.cu:1590-1643), contact rows at their offset ---
  void ccpViolation(double out[4]) {
    out[0] = out[1] = out[2] = out[3] = 0;
    if (!nc) return;
    const int off = (int)bilA.size();
    std::vector<double> dv;
    D.mul(v, dv);
    double mn = 1e300, mv = 1e300, comp = 0, cone = 0;
    for (int c = 0; c < nc; ++c) {
      double gn = gamma[3 * c + off];
      double vn = dv[3 * c + off] + b[3 * c + off];
      mn = std::min(mn, gn);
      mv = std::min(mv, vn);
      comp += gn * vn;
      double gt = std::sqrt(gamma[3 * c + 1 + off] * gamma[3 * c + 1 + off] +
                            gamma[3 * c + 2 + off] * gamma[3 * c + 2 + off]);
      double cv = friction[c] * gn - gt;
      if (cv > 0) cv = 0;
      cone += cv * cv;
    }
    out[0] = std::min(mn, 0.0);
    out[1] = std::min(mv, 0.0);
    out[2] = comp;
    out[3] = std::sqrt(cone);
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// C interface used by the tests (ctypes) and by bench.py's CPU baseline.
extern "C" {

void* or_create() { return new System(); }
void or_destroy(void* s) { delete static_cast<System*>(s); }

void or_set_geom_mode(int mode) { g_geomMode = mode; }
int or_get_geom_mode() { return g_geomMode; }

int or_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

// params: h, mu, tol, maxIt, gx, gy, gz, envelope, bins, warm, antirelax
void or_set_params(void* sp, const double* a) {
  System* s = static_cast<System*>(sp);
  s->h = a[0]; s->mu = a[1]; s->tol = a[2]; s->maxIt = (int)a[3];
  s->gravity = mk(a[4], a[5], a[6]);
  s->envelope = a[7];
  s->bins[0] = s->bins[1] = s->bins[2] = (unsigned)a[8];
  s->warm = a[9] != 0;
  s->antiRelax = a[10] != 0;
}

// solver of the step: 1 APGD (default), 5 PGJ, 6 PGS (System.cu:77-101)
int or_set_solver(void* sp, int kind) {
  if (kind != 1 && kind != 5 && kind != 6) return -1;
  static_cast<System*>(sp)->solverKind = kind;
  return 0;
}

// APGD precision study: 0 FP64, 1 binary32 products, 2 binary32 products and iterates
int or_set_storage(void* sp, int mode) {
  if (mode < 0 || mode > 2) return -1;
  static_cast<System*>(sp)->storage = mode;
  return 0;
}

void or_set_bins(void* sp, int bx, int by, int bz) {
  System* s = static_cast<System*>(sp);
  s->bins[0] = bx; s->bins[1] = by; s->bins[2] = bz;
}

int or_add_bodies(void* sp, int n, const double* pos, const double* vel, const double* mass,
                  const double* geom, const int* fixed, const int* family) {
  System* s = static_cast<System*>(sp);
  for (int i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) {
      s->p.push_back(pos[3 * i + c]);
      s->v.push_back(vel[3 * i + c]);
      s->geom.push_back(geom[3 * i + c]);
    }
    s->bodyMass.push_back(mass[i]);
    s->fixed.push_back(fixed[i]);
    s->family.push_back(family[i]);
  }
  return (int)s->fixed.size();
}

// An ANCF shell mesh as System::importMesh reads it (System.cu:1932-2060):
// nodes (9 per node), connectivity (4 per element), geometry (a, b,
// thickness, contact points per side, 4 per element), external forces (9
// per node), and the M and M^-1 triplets of the shell DOFs (0-based).  Before
// or_initialize, after every body.
int or_add_shell(void* sp, int nNodes, const double* nodes, int nEl, const int* conn, const double* geom,
                 const double* fext, long nM, const int* mI, const int* mJ, const double* mV, long nInv,
                 const int* iI, const int* iJ, const double* iV) {
  System* s = static_cast<System*>(sp);
  if (s->nNodes) return -1;
  s->nNodes = nNodes;
  s->shellNodes.assign(nodes, nodes + 9 * (size_t)nNodes);
  s->shellConn.assign(conn, conn + 4 * (size_t)nEl);
  s->shellGeom.assign(geom, geom + 4 * (size_t)nEl);
  s->shellFext.assign(fext, fext + 9 * (size_t)nNodes);
  std::vector<size_t> perm((size_t)nM);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](size_t a, size_t b) { return mI[a] < mI[b]; });
  s->shellMass = Coo();
  for (size_t q : perm) {
    s->shellMass.I.push_back(mI[q]);
    s->shellMass.J.push_back(mJ[q]);
    s->shellMass.V.push_back(mV[q]);
  }
  s->shellInvI.assign(iI, iI + nInv);
  s->shellInvJ.assign(iJ, iJ + nInv);
  s->shellInvV.assign(iV, iV + nInv);
  return 0;
}

int or_num_geometries(void* sp) { return static_cast<System*>(sp)->ng; }

// Position and shape functions of shell contact point j (geometry nb + j).
void or_shell_point(void* sp, int j, double* pos, double* S) {
  System* s = static_cast<System*>(sp);
  V3 x = s->pointPos(j);
  pos[0] = x.x; pos[1] = x.y; pos[2] = x.z;
  if (S) s->pointShape(j, S);
}

void or_shell_shape(double xi, double eta, double a, double b, double* S) { shellShape(xi, eta, a, b, S); }

int or_add_bilateral(void* sp, int dofA, int dofB, double vel, double t0) {
  System* s = static_cast<System*>(sp);
  s->bilA.push_back(dofA);
  s->bilB.push_back(dofB);
  s->bilVel.push_back(vel);
  s->bilT0.push_back(t0);
  return (int)s->bilA.size();
}

int or_initialize(void* sp) {
  static_cast<System*>(sp)->initialize();
  return 0;
}

int or_step(void* sp) {
  static_cast<System*>(sp)->step();
  return 0;
}

// Run collision detection only (for pair-set checks).
int or_collide(void* sp) {
  System* s = static_cast<System*>(sp);
  s->collide();
  return s->nc;
}

int or_num_contacts(void* sp) { return static_cast<System*>(sp)->nc; }
long or_num_possible(void* sp) { return (long)static_cast<System*>(sp)->numPossible; }
int or_num_rows(void* sp) { return static_cast<System*>(sp)->rows(); }
int or_iterations(void* sp) { return static_cast<System*>(sp)->iterations; }
int or_backtracks(void* sp) { return static_cast<System*>(sp)->backtracks; }
double or_residual(void* sp) { return static_cast<System*>(sp)->residual; }
double or_objective(void* sp) { return static_cast<System*>(sp)->objective; }
double or_time(void* sp) { return static_cast<System*>(sp)->time; }

void or_get_contacts(void* sp, uint32_t* a, uint32_t* b, double* nap, double* loc) {
  System* s = static_cast<System*>(sp);
  for (int c = 0; c < s->nc; ++c) {
    a[c] = s->idA[c];
    b[c] = s->idB[c];
    for (int q = 0; q < 4; ++q) nap[4 * c + q] = s->nAndP[4 * c + q];
    if (loc)
      for (int q = 0; q < 3; ++q) loc[3 * c + q] = s->loc[3 * c + q];
  }
}

// which: 0 p, 1 v, 2 gamma, 3 f_contact, 4 r, 5 b, 6 k, 7 friction
int or_get_vec(void* sp, int which, double* out, long cap) {
  System* s = static_cast<System*>(sp);
  const std::vector<double>* v = nullptr;
  switch (which) {
    case 0: v = &s->p; break;
    case 1: v = &s->v; break;
    case 2: v = &s->gamma; break;
    case 3: v = &s->fContact; break;
    case 4: v = &s->r; break;
    case 5: v = &s->b; break;
    case 6: v = &s->k; break;
    case 7: v = &s->friction; break;
    default: return -1;
  }
  long n = (long)v->size();
  if (out) std::memcpy(out, v->data(), sizeof(double) * (size_t)std::min(n, cap));
  return (int)n;
}

int or_set_vec(void* sp, int which, const double* in, long n) {
  System* s = static_cast<System*>(sp);
  std::vector<double>* v = nullptr;
  switch (which) {
    case 0: v = &s->p; break;
    case 1: v = &s->v; break;
    case 2: v = &s->gamma; break;
    default: return -1;
  }
  v->assign(in, in + n);
  return 0;
}

// D as COO (rows, cols, nnz returned; arrays filled when non-null)
long or_get_D(void* sp, int* I, int* J, double* V, int* rows, int* cols) {
  System* s = static_cast<System*>(sp);
  *rows = s->D.rows;
  *cols = s->D.cols;
  if (I) {
    std::memcpy(I, s->D.I.data(), sizeof(int) * s->D.I.size());
    std::memcpy(J, s->D.J.data(), sizeof(int) * s->D.J.size());
    std::memcpy(V, s->D.V.data(), sizeof(double) * s->D.V.size());
  }
  return (long)s->D.I.size();
}

// y = N x with the current D, M^-1 (APGD::performSchurComplementProduct)
int or_schur(void* sp, const double* x, double* y) {
  System* s = static_cast<System*>(sp);
  std::vector<double> xi(x, x + s->rows()), yo;
  s->schur(xi, yo);
  std::memcpy(y, yo.data(), sizeof(double) * yo.size());
  return 0;
}

// Build k, D, r for the current contacts without solving (for product checks).
int or_assemble(void* sp) {
  System* s = static_cast<System*>(sp);
  s->collide();
  s->buildK();
  s->buildD();
  s->buildR();
  return s->nc;
}

// getResidual (APGD.cu:90-102) of an arbitrary gamma on the assembled system.
double or_residual_of(void* sp, const double* g) {
  System* s = static_cast<System*>(sp);
  std::vector<double> x(g, g + s->rows()), t;
  return s->getResidual(x, t);
}

// The per-row vector whose infinity norm getResidual returns.
void or_residual_vector(void* sp, const double* g, double* out) {
  System* s = static_cast<System*>(sp);
  std::vector<double> x(g, g + s->rows()), t;
  s->getResidual(x, t);
  std::memcpy(out, t.data(), sizeof(double) * t.size());
}

// Cone projection of one 3-vector (APGD.cu:47-78).
void or_project3(double* g, double mu) { System::projectOne(g, mu); }

// Contact tangent frame from a normal (System.cu:753-762).
void or_frame(const double* n, double* u, double* w) {
  V3 uu, ww;
  System::frame(n, &uu, &ww);
  u[0] = uu.x; u[1] = uu.y; u[2] = uu.z;
  w[0] = ww.x; w[1] = ww.y; w[2] = ww.z;
}

// Narrowphase on one pair of bodies: returns 1 if kept, with the stored
// normal (B to A) and gap.
int or_narrow_pair(void* sp, int a, int b, double* n, double* gap) {
  System* s = static_cast<System*>(sp);
  V3 l;
  return s->narrow(a, b, n, gap, &l) ? 1 : 0;
}

void or_ccp_violation(void* sp, double* out) { static_cast<System*>(sp)->ccpViolation(out); }

}  // extern "C"
