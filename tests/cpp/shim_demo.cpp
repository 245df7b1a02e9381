// Drives the C++ shim the way Domain::run drives ParticleManager::step
// (Domain.cpp:139): two overlapping rods, one hard step, prints the results.
#include <cmath>
#include <cstdio>

#include "pcc_particle_manager.hpp"

int main() {
  pcc_params p;
  pcc_default_params(&p);
  p.dt = 1e-3;
  p.tau = 1.0 / std::log(2.0);
  p.bbpgd_tol = p.relcp_tol = p.cutoff_phaseK = 1e-6;
  try {
    pcc::ParticleManager pm(p, 0);
    pm.local_particles.resize(2);
    pm.local_particles[0].gID = 0;
    pm.local_particles[1].gID = 1;
    pm.local_particles[1].position = {0.9, 0.0, 0.0};
    int ghost_calls = 0;
    auto sol = pm.step(0, [&]() { ++ghost_calls; });
    std::printf("constraints %zu relcp %d bbpgd %zu ghost_calls %d\n", sol.constraints.size(),
                sol.constraint_iterations, sol.bbpgd_iterations, ghost_calls);
    for (auto& c : sol.constraints) std::printf("gamma %.17g phi %.17g\n", c.gamma, c.signed_distance);
    for (auto& q : pm.local_particles)
      std::printf("x %.17g length %.17g\n", q.position[0], q.length);
    // failure path: the reference's exception message
    p.max_bbpgd_iterations = 5;
    pcc::ParticleManager bad(p, 0);
    bad.local_particles = pm.local_particles;
    bad.local_particles[1].position = {0.9, 0.0, 0.0};
    try {
      bad.step(1, nullptr);
    } catch (const std::runtime_error& e) {
      std::printf("error: %s\n", e.what());
    }
  } catch (const std::exception& e) {
    std::printf("fatal: %s\n", e.what());
    return 1;
  }
  return 0;
}
