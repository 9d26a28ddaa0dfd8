// C++ drop-in surface of the DVIDE contact time step: the reference's
// System / Body / Solver / APGD / CollisionDetector classes (System.cuh:279-326,
// Body.cuh:13-99, Solver.cuh:17-46, APGD.cuh:18-67, CollisionDetector.cuh:48-103)
// as a header-only layer over the C ABI of libdvide_hip.so (dvide.h).
//
// Drivers keep the reference's shape: new Body, setters, sys->add(body),
// sys->initializeSystem(), loops of sys->DoTimeStep() reading the solver's
// public fields.  Differences a caller sees:
//   * the public device vectors (p_d, v_d, gamma_d, f_contact_d) become
//     host copies returned by p(), v(), gamma(), f_contact(), with setters
//     for p and v;
//   * failures raise dvide::Error (the reference returns 0 and carries on);
//   * System(int solverType) knows the reference's numbering (System.cu:77-101):
//     1 and the default case are APGD; 2..7 (PDIP .. PJKIP) are not on this
//     path and throw dvide::Error(DV_ERR_UNSUPPORTED);
//   * APGD has three builds, chosen by APGD::setBuild (DV_APGD_FP64, the
//     default; DV_APGD_MIXED; DV_APGD_PLAIN); anti-relaxation runs on the
//     FP64 builds (the plain one solves such a scene).
#ifndef DVIDE_SYSTEM_HPP_
#define DVIDE_SYSTEM_HPP_

#include <cmath>
#include <cstdio>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../dvide.h"

namespace dvide {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

// x, y, z triple; converts from any type with x, y, z members (HIP's double3)
struct Vec3 {
  double x = 0, y = 0, z = 0;
  Vec3() = default;
  Vec3(double a, double b, double c) : x(a), y(b), z(c) {}
  template <class T>
  Vec3(const T& t) : x(t.x), y(t.y), z(t.z) {}
};
struct Vec4 {
  double x = 0, y = 0, z = 0, w = 0;
};
inline Vec3 make_vec3(double x, double y, double z) { return Vec3(x, y, z); }

// Body.cuh:13-99 with PhysicsItem's collision family (PhysicsItem.cuh:14-35)
class Body {
 public:
  Body() = default;
  explicit Body(Vec3 position) : pos_(position) {}
  bool isFixed() const { return fixed_; }
  Vec3 getPosition() const { return pos_; }
  Vec3 getVelocity() const { return vel_; }
  double getMass() const { return mass_; }
  void setMass(double m) { mass_ = m; }
  void setBodyFixed(bool f) { fixed_ = f; }
  Vec3 getGeometry() const { return geometry_; }
  void setGeometry(Vec3 g) { geometry_ = g; }
  void setVelocity(Vec3 v) { vel_ = v; }
  int getCollisionFamily() const { return family_; }
  void setCollisionFamily(int f) { family_ = f; }
  unsigned getIndex() const { return index_; }
  unsigned getIdentifier() const { return identifier_; }

 private:
  friend class System;
  Vec3 pos_, vel_;
  double mass_ = 1.0;
  bool fixed_ = false;
  Vec3 geometry_{1.0, 0, 0};
  int family_ = -1;
  unsigned index_ = 0, identifier_ = 0;
};

// Solver.cuh:17-46: public fields read at each DoTimeStep
class Solver {
 public:
  double tolerance = 1e-4;  // APGD's default (APGD.cu:10)
  int maxIterations = 100000;
  int iterations = 0;
  bool verbose = false;
  virtual ~Solver() = default;
};

// APGD.cuh:62-63, plus the choice of build (no counterpart in the
// reference, whose APGD is FP64 only)
class APGD : public Solver {
 public:
  void setWarmStarting(bool b) { warm = b; }
  void setAntiRelaxation(bool b) { antiRelax = b; }
  void setBuild(int b) { build = b; }
  bool warm = false, antiRelax = false;
  int build = DV_APGD_FP64;
};

// CollisionDetector.cuh:89-101
class CollisionDetector {
 public:
  template <class U>
  void setBinsPerAxis(const U& b) { bins[0] = (int)b.x; bins[1] = (int)b.y; bins[2] = (int)b.z; }
  void setEnvelope(double e) { envelope = e; }
  int numCollisions = 0;
  int bins[3] = {20, 20, 20};
  double envelope = 1e-3;  // CollisionDetector.cu:514
};

class System {
 public:
  System() : System(DV_SOLVER_APGD) {}
  explicit System(int solverType) {
    const int rc = dv_scene_create(&s_, solverType);
    if (rc != DV_OK) throw Error(rc, dv_last_error(nullptr));
    solver = &apgd_;
    collisionDetector = &cd_;
  }
  ~System() {
    if (s_) dv_scene_destroy(s_);
  }
  System(const System&) = delete;
  System& operator=(const System&) = delete;

  // --- public state of the reference's System
  std::vector<Body*> bodies;
  Solver* solver;
  CollisionDetector* collisionDetector;
  Vec3 gravity{0, -9.81, 0};  // System.cu:16
  double h = 1e-3, time = 0, elapsedTime = 0, mu = 0.25;
  int timeIndex = 0;

  double getCurrentTime() const { return time; }
  double getTimeStep() const { return h; }
  int getTimeIndex() const { return timeIndex; }
  void setTimeStep(double step) { h = step; }
  void setFrictionCoefficient(double m) { mu = m; }

  int add(Body* b) {
    b->index_ = b->identifier_ = (unsigned)bodies.size();
    bodies.push_back(b);
    return (int)bodies.size();
  }
  // System::importMesh (System.cu:1932-2060): an ANCF shell mesh file, read
  // at initializeSystem after the bodies (its DOFs follow theirs).  Fixed
  // shells only: a mesh with inverse mass makes initializeSystem throw
  // dvide::Error with DV_ERR_UNSUPPORTED.
  void importMesh(const std::string& filename, double stiffness, int numContactPointsPerElement,
                  double pressure = 0.0) {
    meshFile_ = filename;
    meshStiffness_ = stiffness;
    meshPoints_ = numContactPointsPerElement;
    meshPressure_ = pressure;
  }

  int addBilateralConstraintDOF(int a, int b) { return addBilateralConstraintDOF(a, b, 0.0, 0.0); }
  int addBilateralConstraintDOF(int a, int b, double velocity, double startTime) {
    check(dv_add_bilateral_dof(s_, a, b, velocity, startTime));
    return 0;
  }

  int initializeSystem() {
    const size_t n = bodies.size();
    std::vector<double> pos(3 * n), vel(3 * n), mass(n), geom(3 * n);
    std::vector<int> fixed(n), family(n);
    for (size_t i = 0; i < n; ++i) {
      const Body& b = *bodies[i];
      pos[3 * i] = b.pos_.x; pos[3 * i + 1] = b.pos_.y; pos[3 * i + 2] = b.pos_.z;
      vel[3 * i] = b.vel_.x; vel[3 * i + 1] = b.vel_.y; vel[3 * i + 2] = b.vel_.z;
      geom[3 * i] = b.geometry_.x; geom[3 * i + 1] = b.geometry_.y; geom[3 * i + 2] = b.geometry_.z;
      mass[i] = b.mass_;
      fixed[i] = b.fixed_ ? 1 : 0;
      family[i] = b.family_;
    }
    if (n) check(dv_add_bodies(s_, (int)n, pos.data(), vel.data(), mass.data(), geom.data(), fixed.data(),
                               family.data()));
    if (!meshFile_.empty()) check(dv_import_mesh(s_, meshFile_.c_str(), meshStiffness_, meshPoints_, meshPressure_));
    // gravity enters the applied forces once, here (System.cu:344-424)
    check(dv_set_param(s_, DV_PARAM_GX, gravity.x));
    check(dv_set_param(s_, DV_PARAM_GY, gravity.y));
    check(dv_set_param(s_, DV_PARAM_GZ, gravity.z));
    pushParams();
    check(dv_initialize(s_));
    return 0;
  }

  int DoTimeStep() {
    pushParams();
    check(dv_step(s_));
    dv_stats st;
    check(dv_get_stats(s_, &st));
    solver->iterations = st.iterations;
    collisionDetector->numCollisions = st.num_contacts;
    elapsedTime = st.ms_total;
    time = st.time;
    ++timeIndex;
    return 0;
  }

  Vec4 getCCPViolation() {
    double o[4];
    check(dv_ccp_violation(s_, o));
    return Vec4{o[0], o[1], o[2], o[3]};
  }
  double getPotentialEnergy() { return energies()[0]; }
  double getKineticEnergy() { return energies()[1]; }
  double getStrainEnergy() { return energies()[2]; }
  double getTotalEnergy() {
    std::vector<double> e = energies();
    return e[0] + e[1] + e[2];
  }

  // host copies of the reference's public device vectors
  std::vector<double> p() { return vec(DV_VEC_P); }
  std::vector<double> v() { return vec(DV_VEC_V); }
  std::vector<double> gamma() { return vec(DV_VEC_GAMMA); }
  std::vector<double> f_contact() { return vec(DV_VEC_F_CONTACT); }
  void setP(const std::vector<double>& x) { check(dv_set_vector(s_, DV_VEC_P, x.data(), x.size())); }
  void setV(const std::vector<double>& x) { check(dv_set_vector(s_, DV_VEC_V, x.data(), x.size())); }

  // System::exportSystem (System.cu:1699-1735): rigid bodies, default stream
  // formatting as the reference writes them
  int exportSystem(const std::string& filename) {
    std::vector<double> P = p(), V = v();
    std::ofstream f(filename.c_str());
    f << bodies.size() << ", " << 0 << ", " << 0 << ", " << 0 << ", " << std::endl;
    for (size_t i = 0; i < bodies.size(); ++i) {
      const Body& b = *bodies[i];
      f << i << ", " << b.fixed_ << ", " << P[3 * i] << ", " << P[3 * i + 1] << ", " << P[3 * i + 2] << ", "
        << "1, 0, 0, 0, " << V[3 * i] << ", " << V[3 * i + 1] << ", " << V[3 * i + 2] << ", ";
      if (b.geometry_.y == 0) f << "0, " << b.geometry_.x << ", ";
      else f << "2, " << b.geometry_.x << ", " << b.geometry_.y << ", " << b.geometry_.z << ", ";
      f << "\n";
    }
    return 0;
  }

  // System::importSystem (System.cu:1855-1910): adds the file's rigid bodies
  // (masses 2600 * 4 * 3.14159 * r^3 / 3 for a sphere, 1 for a box), owned
  // by this System
  int importSystem(const std::string& filename) {
    std::ifstream f(filename.c_str());
    if (!f) throw Error(DV_ERR_ARG, "cannot open " + filename);
    std::string line;
    std::getline(f, line);
    for (char& c : line)
      if (c == ',') c = ' ';
    int numBodies = 0;
    std::istringstream(line) >> numBodies;
    for (int i = 0; i < numBodies; ++i) {
      if (!std::getline(f, line)) throw Error(DV_ERR_ARG, "truncated system file " + filename);
      for (char& c : line)
        if (c == ',') c = ' ';
      std::istringstream ss(line);
      int index = 0, isFixed = 0, shape = 0;
      double blah = 0;
      Vec3 pos, vel, geometry;
      ss >> index >> isFixed >> pos.x >> pos.y >> pos.z >> blah >> blah >> blah >> blah >> vel.x >> vel.y >>
          vel.z >> shape;
      if (shape == 0) ss >> geometry.x;
      else ss >> geometry.x >> geometry.y >> geometry.z;
      owned_.emplace_back(new Body(pos));
      Body* b = owned_.back().get();
      b->setBodyFixed(isFixed != 0);
      b->setGeometry(geometry);
      b->setVelocity(vel);
      b->setMass(shape == 0 ? 2600 * 4.0 * 3.14159 * std::pow(geometry.x, 3.0) / 3.0 : 1.0);
      add(b);
    }
    return 0;
  }

  // System::exportMatrices (System.cu:1912-1930): D, Minv, r, b and k of the
  // last step as Matrix Market files in `directory` (coordinate 1-based for
  // D and Minv, dense columns for the vectors, 17 significant digits)
  int exportMatrices(const std::string& directory) {
    size_t nnz = 0;
    int rows = 0, cols = 0;
    check(dv_get_jacobian(s_, nullptr, nullptr, nullptr, 0, &nnz, &rows, &cols));
    std::vector<int> I(nnz ? nnz : 1), J(nnz ? nnz : 1);
    std::vector<double> V(nnz ? nnz : 1);
    if (nnz) check(dv_get_jacobian(s_, I.data(), J.data(), V.data(), nnz, &nnz, &rows, &cols));
    writeCoo(directory + "/D.mtx", I, J, V, nnz, rows, cols);
    int nb = 0;
    check(dv_num_bodies(s_, &nb));
    std::vector<double> mass(nb ? nb : 1);
    std::vector<int> fixed(nb ? nb : 1);
    check(dv_get_bodies(s_, mass.data(), nullptr, fixed.data(), nullptr, (size_t)nb));
    const size_t n3 = 3 * (size_t)nb;
    std::vector<int> di(n3 ? n3 : 1);
    std::vector<double> dm(n3 ? n3 : 1);
    for (size_t q = 0; q < n3; ++q) {
      di[q] = (int)q;
      dm[q] = fixed[q / 3] ? 0.0 : 1.0 / mass[q / 3];  // System.cu:380-386
    }
    writeCoo(directory + "/Minv.mtx", di, di, dm, n3, (int)n3, (int)n3);
    writeColumn(directory + "/r.mtx", vec(DV_VEC_R));
    writeColumn(directory + "/b.mtx", vec(DV_VEC_B));
    writeColumn(directory + "/k.mtx", vec(DV_VEC_K));
    return 0;
  }

  dv_scene* handle() { return s_; }

 private:
  void check(int rc) {
    if (rc == DV_OK) return;
    const char* m = dv_last_error(s_);
    throw Error(rc, m ? m : "dvide error");
  }
  static void writeCoo(const std::string& fn, const std::vector<int>& I, const std::vector<int>& J,
                       const std::vector<double>& V, size_t nnz, int rows, int cols) {
    FILE* f = std::fopen(fn.c_str(), "w");
    if (!f) throw Error(DV_ERR_ARG, "cannot write " + fn);
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%d %d %zu\n", rows, cols, nnz);
    for (size_t q = 0; q < nnz; ++q) std::fprintf(f, "%d %d %.17g\n", I[q] + 1, J[q] + 1, V[q]);
    std::fclose(f);
  }
  static void writeColumn(const std::string& fn, const std::vector<double>& x) {
    FILE* f = std::fopen(fn.c_str(), "w");
    if (!f) throw Error(DV_ERR_ARG, "cannot write " + fn);
    std::fprintf(f, "%%%%MatrixMarket matrix array real general\n%zu 1\n", x.size());
    for (double v : x) std::fprintf(f, "%.17g\n", v);
    std::fclose(f);
  }
  void pushParams() {
    check(dv_set_param(s_, DV_PARAM_APGD_BUILD, apgd_.build));
    check(dv_set_param(s_, DV_PARAM_ANTI_RELAX, apgd_.antiRelax ? 1 : 0));
    check(dv_set_param(s_, DV_PARAM_H, h));
    check(dv_set_param(s_, DV_PARAM_MU, mu));
    check(dv_set_param(s_, DV_PARAM_TOL, solver->tolerance));
    check(dv_set_param(s_, DV_PARAM_MAX_ITER, solver->maxIterations));
    check(dv_set_param(s_, DV_PARAM_VERBOSE, solver->verbose ? 1 : 0));
    check(dv_set_param(s_, DV_PARAM_WARM_START, apgd_.warm ? 1 : 0));
    check(dv_set_param(s_, DV_PARAM_ENVELOPE, cd_.envelope));
    check(dv_set_bins_per_axis(s_, cd_.bins[0], cd_.bins[1], cd_.bins[2]));
    check(dv_set_param(s_, DV_PARAM_TIME, time));
  }
  std::vector<double> vec(int which) {
    size_t len = 0;
    check(dv_get_vector(s_, which, nullptr, 0, &len));
    std::vector<double> out(len);
    if (len) check(dv_get_vector(s_, which, out.data(), len, &len));
    return out;
  }
  std::vector<double> energies() {
    std::vector<double> e(3);
    check(dv_energies(s_, e.data()));
    return e;
  }

  dv_scene* s_ = nullptr;
  std::vector<std::unique_ptr<Body>> owned_;  // bodies created by importSystem
  std::string meshFile_;                      // importMesh's arguments
  double meshStiffness_ = 0, meshPressure_ = 0;
  int meshPoints_ = 0;
  APGD apgd_;
  CollisionDetector cd_;
};

}  // namespace dvide

#endif  // DVIDE_SYSTEM_HPP_
