// A reference-style driver (models/test_oneBall.cu:228-262) on the C++
// drop-in surface: a ball resting on a fixed ground box, stepped, with its
// normal impulse, contact force and CCP violation printed for the test to
// check (gamma_n = m g h, f_contact = m g); then System::exportSystem,
// System::exportMatrices into argv[2], a second System built by
// System::importSystem from the exported file and stepped once, the
// factory's answer to a solver type that is not on this path (2 = PDIP),
// and, given a mesh file and a position, a ball resting on a fixed shell
// read by System::importMesh.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "dvide/System.hpp"

using dvide::APGD;
using dvide::Body;
using dvide::System;
using dvide::make_vec3;

int main(int argc, char** argv) {
  const int build = argc > 1 ? std::atoi(argv[1]) : DV_APGD_FP64;
  const std::string dir = argc > 2 ? argv[2] : ".";
  try {
    System* sys = new System(1);
    dynamic_cast<APGD*>(sys->solver)->setBuild(build);
    sys->setTimeStep(1e-2);
    sys->solver->tolerance = 1e-8;
    sys->solver->maxIterations = 10000;
    Body* ground = new Body(make_vec3(0, -0.5, 0));
    ground->setBodyFixed(true);
    ground->setGeometry(make_vec3(5, 0.5, 5));
    sys->add(ground);
    Body* ball = new Body(make_vec3(0, 0.5, 0));
    ball->setGeometry(make_vec3(0.5, 0, 0));
    ball->setMass(1.0);
    sys->add(ball);
    sys->initializeSystem();
    for (int k = 0; k < 3; ++k) sys->DoTimeStep();
    std::vector<double> g = sys->gamma(), f = sys->f_contact(), v = sys->v();
    dvide::Vec4 viol = sys->getCCPViolation();
    std::printf("collisions %d iterations %d gamma %.17g %.17g %.17g f_contact_y %.17g v_y %.17g time %.17g "
                "violation_x %.17g violation_y %.17g violation_z %.17g violation_w %.17g\n",
                sys->collisionDetector->numCollisions, sys->solver->iterations, g[0], g[1], g[2], f[4], v[4],
                sys->getCurrentTime(), viol.x, viol.y, viol.z, viol.w);
    sys->exportSystem(dir + "/system.csv");
    sys->exportMatrices(dir);
    delete sys;
    delete ball;
    delete ground;

    System* again = new System();
    dynamic_cast<APGD*>(again->solver)->setBuild(build);
    again->setTimeStep(1e-2);
    again->solver->tolerance = 1e-8;
    again->solver->maxIterations = 10000;
    again->importSystem(dir + "/system.csv");
    again->initializeSystem();
    again->DoTimeStep();
    std::vector<double> g2 = again->gamma();
    std::printf("imported_bodies %zu imported_mass %.17g imported_collisions %d imported_gamma %.17g\n",
                again->bodies.size(), again->bodies[1]->getMass(), again->collisionDetector->numCollisions,
                g2[0]);
    delete again;
  } catch (const dvide::Error& e) {
    std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
    return 2;
  }
  if (argc > 7) {
    // a ball resting on a contact point of the fixed shell mesh in argv[3]
    // (System::importMesh), centred at argv[4..6] with radius argv[7]
    try {
      System* s3 = new System(1);
      dynamic_cast<APGD*>(s3->solver)->setBuild(build);
      s3->setTimeStep(1e-3);
      s3->solver->tolerance = 1e-8;
      s3->solver->maxIterations = 10000;
      Body* ball = new Body(make_vec3(std::atof(argv[4]), std::atof(argv[5]), std::atof(argv[6])));
      ball->setGeometry(make_vec3(std::atof(argv[7]), 0, 0));
      ball->setMass(1.0);
      s3->add(ball);
      s3->importMesh(argv[3], 2e8, 8, 0.0);
      s3->initializeSystem();
      s3->DoTimeStep();
      std::vector<double> g3 = s3->gamma(), p3 = s3->p();
      std::printf("shell_collisions %d shell_gamma %.17g shell_dofs %zu\n", s3->collisionDetector->numCollisions,
                  g3[0], p3.size());
      delete s3;
      delete ball;
    } catch (const dvide::Error& e) {
      std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
      return 3;
    }
  }
  try {
    System pdip(2);
    std::printf("factory_pdip accepted\n");
  } catch (const dvide::Error& e) {
    std::printf("factory_pdip %d\n", e.code);
  }
  return 0;
}
