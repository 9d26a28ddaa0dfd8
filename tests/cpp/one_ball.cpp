// A reference-style driver (models/test_oneBall.cu:228-262) on the C++
// drop-in surface: a ball resting on a fixed ground box, stepped, with its
// normal impulse and contact force printed for the test to check
// (gamma_n = m g h, f_contact = m g).
#include <cstdio>

#include "dvide/System.hpp"

using dvide::Body;
using dvide::System;
using dvide::make_vec3;

int main(int argc, char** argv) {
  const int solverType = argc > 1 ? std::atoi(argv[1]) : DV_SOLVER_APGD_FUSED;
  try {
    System* sys = new System(solverType);
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
                "violation %.3g\n",
                sys->collisionDetector->numCollisions, sys->solver->iterations, g[0], g[1], g[2], f[4], v[4],
                sys->getCurrentTime(), viol.w);
    delete sys;
    delete ball;
    delete ground;
  } catch (const dvide::Error& e) {
    std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
    return 2;
  }
  return 0;
}
