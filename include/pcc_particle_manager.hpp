// pcc_particle_manager.hpp — header-only C++ shim over the C ABI (pcc.h) that
// keeps the reference's step interface:
//
//   ParticleManager::SolverSolution ParticleManager::step(int i,
//       std::function<void()> exchangeGhostParticles)
//   (code/cpp/src/simulation/ParticleManager.h:50, ParticleManager.cpp:150-166)
//
// `local_particles` is marshalled to the device before the step and back after
// it; the constraint vector of the step is returned in the SolverSolution
// (ParticleManager.h:32-38).  Solver failures throw std::runtime_error with the
// reference's messages (BBPGD.cpp:120, Hard.cpp:189, ParticleManager.cpp:165).
// There is one GPU and no domain decomposition, so the ghost-exchange callback
// is invoked once per step and ghost_particles stays empty.
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "pcc.h"

namespace pcc {

// The fields of Particle/ParticleData (ParticleData.h:8-17, Particle.h:154-165)
// that the step reads or writes.
struct ParticleRecord {
  int32_t gID = 0;
  std::array<double, 3> position{};
  std::array<double, 4> quaternion{1.0, 0.0, 0.0, 0.0};
  double length = 1.0, l0 = 1.0, diameter = 0.5, ldot = 0.0;
  std::array<double, 3> force{}, torque{}, velocityLinear{}, velocityAngular{};
  double impedance = 0.0, stress = 0.0;
  int num_constraints = 0;
  int age = 1;  // generation: 1 for a new cell, +1 for both halves at each division
};

// Constraint (code/cpp/src/dynamics/Constraint.h:6-49), host copy.
struct ConstraintRecord {
  double signed_distance;
  int gidI, gidJ;
  std::array<double, 3> normI, rPosI, rPosJ, contactPoint;
  double stressI, stressJ;
  int gid, iteration;
  double gamma;
};

class ParticleManager {
 public:
  struct SolverSolution {
    std::vector<ConstraintRecord> constraints;
    int constraint_iterations;
    size_t bbpgd_iterations;
    double residual;
    double max_overlap;
  };

  explicit ParticleManager(const pcc_params& params, int device = 0) {
    pcc_status s = pcc_create(&params, device, &ctx_);
    if (s != PCC_OK) throw std::runtime_error(pcc_status_string(s));
  }
  ~ParticleManager() { pcc_destroy(ctx_); }
  ParticleManager(const ParticleManager&) = delete;
  ParticleManager& operator=(const ParticleManager&) = delete;

  SolverSolution step(int i, std::function<void()> exchangeGhostParticles) {
    if (exchangeGhostParticles) exchangeGhostParticles();
    upload();
    pcc_step_result r{};
    pcc_status s = pcc_step(ctx_, i, &r);
    // the reference mutates particles before throwing RELCP_MAX_ITERS
    if (s == PCC_OK || s == PCC_E_RELCP_MAX_ITERS) download();
    if (s != PCC_OK) throw std::runtime_error(pcc_last_error(ctx_));
    SolverSolution out;
    out.constraints = constraints(r.n_constraints);
    out.constraint_iterations = r.constraint_iterations;
    out.bbpgd_iterations = (size_t)r.bbpgd_iterations;
    out.residual = r.residual;
    out.max_overlap = r.max_overlap;
    return out;
  }

  std::vector<ParticleRecord> local_particles;  // gid order
  std::vector<ParticleRecord> ghost_particles;  // always empty on one GPU

 private:
  pcc_ctx* ctx_ = nullptr;
  std::vector<double> pos_, quat_, len_, l0_, diam_, ldot_, f_, t_, v_, w_, imp_, str_;
  std::vector<int32_t> nc_, age_;

  pcc_bodies view(size_t n) {
    pos_.resize(3 * n);
    quat_.resize(4 * n);
    len_.resize(n);
    l0_.resize(n);
    diam_.resize(n);
    ldot_.resize(n);
    f_.resize(3 * n);
    t_.resize(3 * n);
    v_.resize(3 * n);
    w_.resize(3 * n);
    imp_.resize(n);
    str_.resize(n);
    nc_.resize(n);
    age_.resize(n);
    pcc_bodies b{};
    b.n = (int64_t)n;
    b.pos = pos_.data();
    b.quat = quat_.data();
    b.length = len_.data();
    b.l0 = l0_.data();
    b.diameter = diam_.data();
    b.ldot = ldot_.data();
    b.force = f_.data();
    b.torque = t_.data();
    b.velocity = v_.data();
    b.omega = w_.data();
    b.impedance = imp_.data();
    b.stress = str_.data();
    b.num_constraints = nc_.data();
    b.age = age_.data();
    return b;
  }

  void upload() {
    const size_t n = local_particles.size();
    pcc_bodies b = view(n);
    for (size_t i = 0; i < n; ++i) {
      const ParticleRecord& p = local_particles[i];
      if (p.gID != (int32_t)i) throw std::runtime_error("particles must be in gid order 0..N-1");
      for (int k = 0; k < 3; ++k) pos_[3 * i + k] = p.position[k];
      for (int k = 0; k < 4; ++k) quat_[4 * i + k] = p.quaternion[k];
      len_[i] = p.length;
      l0_[i] = p.l0;
      diam_[i] = p.diameter;
      ldot_[i] = p.ldot;
      imp_[i] = p.impedance;  // input too: the Brownian turn scales with it
      age_[i] = p.age;
    }
    pcc_status s = pcc_set_state(ctx_, &b);
    if (s != PCC_OK) throw std::runtime_error(pcc_last_error(ctx_));
  }

  void download() {
    const size_t n = local_particles.size();
    pcc_bodies b = view(n);
    pcc_status s = pcc_get_state(ctx_, &b);
    if (s != PCC_OK) throw std::runtime_error(pcc_last_error(ctx_));
    for (size_t i = 0; i < n; ++i) {
      ParticleRecord& p = local_particles[i];
      for (int k = 0; k < 3; ++k) {
        p.position[k] = pos_[3 * i + k];
        p.force[k] = f_[3 * i + k];
        p.torque[k] = t_[3 * i + k];
        p.velocityLinear[k] = v_[3 * i + k];
        p.velocityAngular[k] = w_[3 * i + k];
      }
      for (int k = 0; k < 4; ++k) p.quaternion[k] = quat_[4 * i + k];
      p.length = len_[i];
      p.ldot = ldot_[i];
      p.impedance = imp_[i];
      p.stress = str_[i];
      p.num_constraints = nc_[i];
      p.age = age_[i];
    }
  }

  std::vector<ConstraintRecord> constraints(int64_t count) {
    std::vector<ConstraintRecord> out((size_t)count);
    if (count == 0) return out;
    std::vector<int32_t> gi(count), gj(count), gid(count), it(count);
    std::vector<double> phi(count), n(3 * count), ri(3 * count), rj(3 * count), cp(3 * count),
        si(count), sj(count), g(count);
    pcc_constraints c{};
    c.capacity = count;
    c.gidI = gi.data();
    c.gidJ = gj.data();
    c.signed_distance = phi.data();
    c.normal = n.data();
    c.rPosI = ri.data();
    c.rPosJ = rj.data();
    c.contact_point = cp.data();
    c.stressI = si.data();
    c.stressJ = sj.data();
    c.gamma = g.data();
    c.gid = gid.data();
    c.iteration = it.data();
    int64_t got = 0;
    pcc_status s = pcc_get_constraints(ctx_, &c, &got);
    if (s != PCC_OK) throw std::runtime_error(pcc_last_error(ctx_));
    for (int64_t i = 0; i < count; ++i) {
      ConstraintRecord& r = out[(size_t)i];
      r.signed_distance = phi[i];
      r.gidI = gi[i];
      r.gidJ = gj[i];
      for (int k = 0; k < 3; ++k) {
        r.normI[k] = n[3 * i + k];
        r.rPosI[k] = ri[3 * i + k];
        r.rPosJ[k] = rj[3 * i + k];
        r.contactPoint[k] = cp[3 * i + k];
      }
      r.stressI = si[i];
      r.stressJ = sj[i];
      r.gid = gid[i];
      r.iteration = it[i];
      r.gamma = g[i];
    }
    return out;
  }
};

}  // namespace pcc
