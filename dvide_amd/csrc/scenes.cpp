// Seeded scene generators for the contact time step.
//
// These play the role of the reference's scenario drivers (models/*.cu), which
// build a scene body by body through System::add(Body*) and then loop
// DoTimeStep.  The reference seeds nothing (C rand(), models/test_draft.cu),
// so every scene here is built from std::mt19937_64 with the explicit 53-bit
// uniform u = (x >> 11) * 2^-53, which is fully specified by the C++ standard
// and therefore reproducible on any host.
//
// Layout conventions follow the reference:
//   * y is up, gravity (0,-9.81,0)                        (System.cu:16)
//   * a Body is a 3-DOF point mass; geometry (r,0,0) is a sphere and
//     (hx,hy,hz) with hy != 0 an axis-aligned box          (Body.cuh:45-98,
//                                                            CollisionDetector.cu:72-75)
//   * container walls/ground/lid are boxes, added before the spheres, as in
//     models/test_compaction.cu:272-315
//   * mass = rho * 4/3 * pi * r^3
//
// Built into libdvide_scenes.so (host only, no HIP), used by the tests, the
// benchmark and the C++ drivers.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Rng {
  std::mt19937_64 g;
  explicit Rng(uint64_t seed) : g(seed) {}
  double u() { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }
  double uni(double a, double b) { return a + (b - a) * u(); }
};

const double kPi = 3.14159265358979323846;

struct Scene {
  std::vector<double> pos, vel, mass, geom;
  std::vector<int> fixed, family;
  std::vector<int> bil_dofs;     // 2 per bilateral constraint (dofA, dofB or -1)
  std::vector<double> bil_info;  // 2 per bilateral constraint (velocity, start time)
  double params[16];
  Scene() {
    for (double& p : params) p = 0;
  }
  int add(double x, double y, double z, double gx, double gy, double gz, double m,
          bool isFixed, int fam = -1) {
    pos.insert(pos.end(), {x, y, z});
    vel.insert(vel.end(), {0.0, 0.0, 0.0});
    geom.insert(geom.end(), {gx, gy, gz});
    mass.push_back(m);
    fixed.push_back(isFixed ? 1 : 0);
    family.push_back(fam);
    return (int)mass.size() - 1;
  }
  int addSphere(double x, double y, double z, double r, double rho) {
    return add(x, y, z, r, 0, 0, rho * 4.0 / 3.0 * kPi * r * r * r, false);
  }
  int addBox(double x, double y, double z, double hx, double hy, double hz, bool isFixed,
             double m = 1.0) {
    return add(x, y, z, hx, hy, hz, m, isFixed);
  }
};

// Parameter slots of Scene::params (mirrored in dvide_amd/scenes.py).
enum {
  P_H = 0, P_MU, P_TOL, P_MAXIT, P_GX, P_GY, P_GZ, P_ENVELOPE, P_BINS, P_STEPS,
  P_WARM, P_NPARAMS
};

void defaultParams(Scene& s) {
  s.params[P_H] = 1e-3;
  s.params[P_MU] = 0.25;
  s.params[P_TOL] = 1e-4;
  s.params[P_MAXIT] = 500;
  s.params[P_GX] = 0;
  s.params[P_GY] = -9.81;
  s.params[P_GZ] = 0;
  s.params[P_ENVELOPE] = 1e-3;
  s.params[P_BINS] = 20;
  s.params[P_STEPS] = 1;
  s.params[P_WARM] = 0;
}

// Open-top container: ground slab plus four walls around the inner region
// [0,W]x[0,H]x[0,D], wall thickness 2*t.  Boxes overlap at the corners, which
// is harmless because box-box pairs never produce contacts
// (CollisionDetector.cu:287-328 handles sphere-sphere and sphere-box only).
void addContainer(Scene& s, double W, double H, double D, double t) {
  s.addBox(0.5 * W, -t, 0.5 * D, 0.5 * W + 2 * t, t, 0.5 * D + 2 * t, true);   // ground
  s.addBox(-t, 0.5 * H, 0.5 * D, t, 0.5 * H + t, 0.5 * D + 2 * t, true);       // x = 0
  s.addBox(W + t, 0.5 * H, 0.5 * D, t, 0.5 * H + t, 0.5 * D + 2 * t, true);    // x = W
  s.addBox(0.5 * W, 0.5 * H, -t, 0.5 * W + 2 * t, 0.5 * H + t, t, true);       // z = 0
  s.addBox(0.5 * W, 0.5 * H, D + t, 0.5 * W + 2 * t, 0.5 * H + t, t, true);    // z = D
}

// Sequential deposition of polydisperse spheres into an open box
// [0,W]x[0,inf)x[0,D].  Spheres are dropped one at a time above a lattice
// column; each descends in small steps and after every step is projected out
// of the spheres already placed, the floor and the walls (Gauss-Seidel over
// its own contacts), which lets it roll into a resting pocket.  A sphere stops
// when a step no longer lowers it.  The pile has no overlaps (up to 1e-12 m)
// and its resting contacts have gaps near zero, i.e. a settled state.
void deposit(Rng& rng, int n, double rmin, double rmax, double W, double D,
             std::vector<double>& x, std::vector<double>& r) {
  r.resize(n);
  x.assign(3 * (size_t)n, 0.0);
  for (int i = 0; i < n; ++i) r[i] = rng.uni(rmin, rmax);
  const double cs = 2.0 * rmax * 1.0001;
  const int nx = std::max(1, (int)std::ceil(W / cs)), nz = std::max(1, (int)std::ceil(D / cs));
  std::vector<double> top((size_t)nx * nz, 0.0);   // column heightmap
  std::vector<std::vector<int>> cells;            // (cy*nz+cz)*nx+cx -> placed spheres
  auto cellIndex = [&](int cx, int cy, int cz) { return ((size_t)cy * nz + cz) * nx + cx; };
  const int ncol = nx * nz;
  std::vector<int> nb;
  for (int i = 0; i < n; ++i) {
    const double ri = r[i];
    int col = i % ncol;
    int ccx = col % nx, ccz = col / nx;
    double px = std::min(W - ri, std::max(ri, (ccx + 0.5) * cs + rng.uni(-0.25, 0.25) * cs));
    double pz = std::min(D - ri, std::max(ri, (ccz + 0.5) * cs + rng.uni(-0.25, 0.25) * cs));
    double h = 0;
    for (int a = -1; a <= 1; ++a)
      for (int b = -1; b <= 1; ++b) {
        int X = ccx + a, Z = ccz + b;
        if (X >= 0 && Z >= 0 && X < nx && Z < nz) h = std::max(h, top[(size_t)Z * nx + X]);
      }
    double py = h + ri + 1e-3;
    for (int step = 0; step < 4000; ++step) {
      const double xPrev = px, yPrev = py, zPrev = pz;
      py -= 0.25 * ri;
      int cx = std::min(nx - 1, std::max(0, (int)std::floor(px / cs)));
      int cz = std::min(nz - 1, std::max(0, (int)std::floor(pz / cs)));
      int cy = std::max(0, (int)std::floor(py / cs));
      nb.clear();
      for (int oy = -2; oy <= 2; ++oy)
        for (int oz = -2; oz <= 2; ++oz)
          for (int ox = -2; ox <= 2; ++ox) {
            int X = cx + ox, Y = cy + oy, Z = cz + oz;
            if (X < 0 || Y < 0 || Z < 0 || X >= nx || Z >= nz) continue;
            size_t c = cellIndex(X, Y, Z);
            if (c >= cells.size()) continue;
            for (int j : cells[c]) nb.push_back(j);
          }
      double worst = 0;
      for (int pass = 0; pass < 200; ++pass) {
        worst = 0;
        for (int j : nb) {
          double ex = px - x[3 * j], ey = py - x[3 * j + 1], ez = pz - x[3 * j + 2];
          double d2 = ex * ex + ey * ey + ez * ez, s = ri + r[j];
          if (d2 >= s * s) continue;
          double d = std::sqrt(d2);
          if (d < 1e-12) { ex = 0; ey = 1; ez = 0; d = 1; }
          double ov = s - d;
          worst = std::max(worst, ov);
          px += ov * ex / d; py += ov * ey / d; pz += ov * ez / d;
        }
        if (py < ri) { worst = std::max(worst, ri - py); py = ri; }
        px = std::min(W - ri, std::max(ri, px));
        pz = std::min(D - ri, std::max(ri, pz));
        if (worst < 1e-12) break;
      }
      if (worst >= 1e-12) {  // wedged: keep the last overlap-free position
        px = xPrev; py = yPrev; pz = zPrev;
        break;
      }
      if (yPrev - py < 1e-7) break;
    }
    x[3 * i] = px; x[3 * i + 1] = py; x[3 * i + 2] = pz;
    int cx = std::min(nx - 1, std::max(0, (int)std::floor(px / cs)));
    int cz = std::min(nz - 1, std::max(0, (int)std::floor(pz / cs)));
    int cy = std::max(0, (int)std::floor(py / cs));
    size_t need = (size_t)(cy + 4) * nz * nx;
    if (cells.size() < need) cells.resize(need);
    cells[cellIndex(cx, cy, cz)].push_back(i);
    double& t = top[(size_t)cz * nx + cx];
    t = std::max(t, py + ri);
  }
}

// --- scene kinds ------------------------------------------------------------

// Ball resting on a fixed ground box (models/test_oneBall.cu:241-247): unit
// mass, radius 0.5, touching the box top.  a0 = lateral gravity g_x.
Scene* ballOnBox(double gx, double h) {
  Scene* s = new Scene();
  defaultParams(*s);
  s->params[P_H] = h > 0 ? h : 1e-2;
  s->params[P_GX] = gx;
  s->params[P_TOL] = 1e-8;
  s->params[P_MAXIT] = 10000;
  s->addBox(0, -0.5, 0, 5, 0.5, 5, true);
  int b = s->add(0, 0.5, 0, 0.5, 0, 0, 1.0, false);
  (void)b;
  return s;
}

// 1,000 spheres r = 0.05, rho = 2500, mu = 0.25 on a 10x10x10 jittered lattice
// (pitch 0.12, jitter U(+-0.01)) in a 1.2 m open box.
Scene* latticeBox(uint64_t seed, int perSide) {
  Scene* s = new Scene();
  defaultParams(*s);
  s->params[P_MU] = 0.25;
  s->params[P_STEPS] = 200;
  s->params[P_MAXIT] = 500;
  const double r = 0.05, pitch = 0.12, W = pitch * perSide;
  addContainer(*s, W, W + 0.2, W, 0.05);
  Rng rng(seed);
  for (int i = 0; i < perSide; ++i)
    for (int j = 0; j < perSide; ++j)
      for (int k = 0; k < perSide; ++k) {
        double x = 0.5 * pitch + i * pitch + rng.uni(-0.01, 0.01);
        double y = 0.5 * pitch + j * pitch + rng.uni(-0.01, 0.01);
        double z = 0.5 * pitch + k * pitch + rng.uni(-0.01, 0.01);
        s->addSphere(x, y, z, r, 2500.0);
      }
  return s;
}

// Granular compaction: n spheres r ~ U[0.04,0.06], mu = 0.3, settled in a
// W x W open box and capped by a rigid lid box moving down at 0.1 m/s.
//   lidMode 0: a free lid as heavy as the spheres together, pressing with its
//     weight (models/test_compaction.cu:272-277), released at -0.1 m/s, with
//     APGD warm starting as that driver sets it (test_compaction.cu:245-247);
//   lidMode 2: no lid (the settled pile under gravity alone);
//   lidMode 1: a 50 kg lid driven at -0.1 m/s through a bilateral DOF row on
//     its y coordinate (System.cu:522-532), its x and z held by two more rows.
//     Driving a rigid lid at a fixed speed into a jammed packing of rigid
//     spheres has no solution once the lid bears on the packing, so APGD
//     stalls from the third step on; this mode serves the bilateral rows'
//     tests on the first steps.
Scene* compaction(uint64_t seed, int n, double W, int sweeps, int lidMode, double lidMass,
                  double lidSpeed) {
  Scene* s = new Scene();
  defaultParams(*s);
  s->params[P_MU] = 0.3;
  s->params[P_MAXIT] = 1000;
  s->params[P_STEPS] = 3;
  s->params[P_WARM] = lidMode == 0 ? 1 : 0;
  Rng rng(seed);
  std::vector<double> x, r;
  deposit(rng, n, 0.04, 0.06, W, W, x, r);
  (void)sweeps;
  double top = 0;
  for (int i = 0; i < n; ++i) top = std::max(top, x[3 * i + 1] + r[i]);
  addContainer(*s, W, top + 0.3, W, 0.05);
  const double lt = 0.05;
  if (lidMode == 2) {
    for (int i = 0; i < n; ++i) s->addSphere(x[3 * i], x[3 * i + 1], x[3 * i + 2], r[i], 2500.0);
    s->params[P_WARM] = 1;
    return s;
  }
  // lid bottom face at the highest sphere top: the top layer is in contact
  int lid = s->addBox(0.5 * W, top + lt, 0.5 * W, 0.5 * W - 0.001, lt, 0.5 * W - 0.001, false,
                      50.0);
  double total = 0;
  for (int i = 0; i < n; ++i) {
    const int b = s->addSphere(x[3 * i], x[3 * i + 1], x[3 * i + 2], r[i], 2500.0);
    total += s->mass[b];
  }
  if (lidMode == 0) {
    s->mass[lid] = lidMass > 0 ? lidMass : total;
    s->vel[3 * lid + 1] = -lidSpeed;
  } else {
    s->bil_dofs.insert(s->bil_dofs.end(), {3 * lid + 0, -1, 3 * lid + 1, -1, 3 * lid + 2, -1});
    s->bil_info.insert(s->bil_info.end(), {0.0, 0.0, -0.1, 0.0, 0.0, 0.0});
  }
  return s;
}

// Sphere pile sized for roughly `targetContacts` contacts, radii U[rmin,rmax],
// mu given; used for the batched and served scene families.
Scene* pile(uint64_t seed, int targetContacts, double rmin, double rmax, double mu,
            int sweeps) {
  Scene* s = new Scene();
  defaultParams(*s);
  s->params[P_MU] = mu;
  s->params[P_MAXIT] = 1000;
  Rng rng(seed);
  int n = std::max(8, (int)std::lround(targetContacts / 3.6));
  // footprint ~ square, pile height ~ footprint width
  double vol = n * 4.0 / 3.0 * kPi * std::pow(0.5 * (rmin + rmax), 3) / 0.58;
  double W = std::max(6 * rmax, std::cbrt(vol));
  std::vector<double> x, r;
  deposit(rng, n, rmin, rmax, W, W, x, r);
  (void)sweeps;
  double top = 0;
  for (int i = 0; i < n; ++i) top = std::max(top, x[3 * i + 1] + r[i]);
  addContainer(*s, W, top + 0.3, W, 0.05);
  for (int i = 0; i < n; ++i) s->addSphere(x[3 * i], x[3 * i + 1], x[3 * i + 2], r[i], 2500.0);
  return s;
}

// Uniform random spheres in a cube (overlaps allowed): a collision-detection
// stress case with a box floor, used for brute-force pair-set checks.
Scene* cloud(uint64_t seed, int n, double L, double rmin, double rmax) {
  Scene* s = new Scene();
  defaultParams(*s);
  Rng rng(seed);
  s->addBox(0.5 * L, -0.05, 0.5 * L, 0.5 * L + 0.1, 0.05, 0.5 * L + 0.1, true);
  for (int i = 0; i < n; ++i) {
    double x = rng.uni(0, L), y = rng.uni(0, L), z = rng.uni(0, L), r = rng.uni(rmin, rmax);
    int b = s->addSphere(x, y, z, r, 2500.0);
    if (i % 7 == 3) s->family[b] = 1 + (i % 3);  // a few collision families
  }
  // a second, free box inside the cloud
  s->addBox(0.3 * L, 0.5 * L, 0.6 * L, 0.1 * L, 0.05 * L, 0.08 * L, false, 5.0);
  return s;
}

// Spheres far apart in free fall (no contacts): System.cu:575-580 branch.
Scene* freeFall(uint64_t seed, int n) {
  Scene* s = new Scene();
  defaultParams(*s);
  Rng rng(seed);
  for (int i = 0; i < n; ++i) {
    int b = s->addSphere(i * 1.0, 10.0 + rng.uni(0, 1), 0, 0.1, 1000.0);
    s->vel[3 * b] = rng.uni(-1, 1);
  }
  return s;
}

}  // namespace

extern "C" {

enum {
  DVS_BALL_ON_BOX = 0,
  DVS_LATTICE_BOX = 1,
  DVS_COMPACTION = 2,
  DVS_PILE = 3,
  DVS_CLOUD = 4,
  DVS_FREE_FALL = 5,
};

// Build a scene.  `a` holds kind-specific arguments:
//   BALL_ON_BOX: a[0] = lateral gravity g_x, a[1] = h
//   LATTICE_BOX: a[0] = spheres per side (default 10)
//   COMPACTION:  a[0] = sphere count, a[1] = footprint width, a[2] = settle sweeps,
//                a[3] = lid mode (0 free heavy lid, 1 bilateral-driven lid),
//                a[4] = free lid mass (0: the spheres' total), a[5] = free lid
//                speed (0: 0.1 m/s, negative: at rest)
//   PILE:        a[0] = target contacts, a[1] = rmin, a[2] = rmax, a[3] = mu, a[4] = sweeps
//   CLOUD:       a[0] = sphere count, a[1] = cube edge, a[2] = rmin, a[3] = rmax
//   FREE_FALL:   a[0] = sphere count
// Returns NULL for an unknown kind.
void* dvs_generate(int kind, uint64_t seed, const double* a) {
  switch (kind) {
    case DVS_BALL_ON_BOX: return ballOnBox(a[0], a[1]);
    case DVS_LATTICE_BOX: return latticeBox(seed, a[0] > 0 ? (int)a[0] : 10);
    case DVS_COMPACTION:
      return compaction(seed, a[0] > 0 ? (int)a[0] : 64000, a[1] > 0 ? a[1] : 4.0,
                        a[2] > 0 ? (int)a[2] : 400, (int)a[3], a[4], a[5] < 0 ? 0.0 : (a[5] > 0 ? a[5] : 0.1));
    case DVS_PILE:
      return pile(seed, a[0] > 0 ? (int)a[0] : 1000, a[1] > 0 ? a[1] : 0.03,
                  a[2] > 0 ? a[2] : 0.06, a[3] > 0 ? a[3] : 0.3, a[4] > 0 ? (int)a[4] : 300);
    case DVS_CLOUD:
      return cloud(seed, a[0] > 0 ? (int)a[0] : 500, a[1] > 0 ? a[1] : 1.0,
                   a[2] > 0 ? a[2] : 0.02, a[3] > 0 ? a[3] : 0.06);
    case DVS_FREE_FALL: return freeFall(seed, a[0] > 0 ? (int)a[0] : 10);
  }
  return nullptr;
}

void dvs_free(void* s) { delete static_cast<Scene*>(s); }
int dvs_num_bodies(const void* s) { return (int)static_cast<const Scene*>(s)->mass.size(); }
int dvs_num_bilaterals(const void* s) {
  return (int)static_cast<const Scene*>(s)->bil_dofs.size() / 2;
}
const double* dvs_pos(const void* s) { return static_cast<const Scene*>(s)->pos.data(); }
const double* dvs_vel(const void* s) { return static_cast<const Scene*>(s)->vel.data(); }
const double* dvs_mass(const void* s) { return static_cast<const Scene*>(s)->mass.data(); }
const double* dvs_geom(const void* s) { return static_cast<const Scene*>(s)->geom.data(); }
const int* dvs_fixed(const void* s) { return static_cast<const Scene*>(s)->fixed.data(); }
const int* dvs_family(const void* s) { return static_cast<const Scene*>(s)->family.data(); }
const int* dvs_bil_dofs(const void* s) { return static_cast<const Scene*>(s)->bil_dofs.data(); }
const double* dvs_bil_info(const void* s) {
  return static_cast<const Scene*>(s)->bil_info.data();
}
const double* dvs_params(const void* s) { return static_cast<const Scene*>(s)->params; }

}  // extern "C"
