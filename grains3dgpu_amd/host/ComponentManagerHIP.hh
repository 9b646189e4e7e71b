// ComponentManagerHIP.hh -- C++ host mirror of the reference's component
// manager, driving libgrains_hip.so through include/grains_hip.h.
//
// The reference's driver (Grains/Base/src/GrainsCPU.cpp:46-70) calls, per
// time step, ComponentManager<T>::detectCollisionAndComputeContactForces,
// addExternalForces and moveParticles (ComponentManager.hh:132-184), and
// reads/writes the state through getTransform/setTransform, getVelocity/
// setVelocity and getTorce (ComponentManager.hh:40-97).  This class keeps
// those method names and their argument meaning with plain C++ types; the
// three per-step calls are fused on the device, so the step is launched by
// moveParticles() (the last of the three in the reference's loop) and the
// two earlier calls only record their inputs.  Errors throw
// std::runtime_error carrying grains_last_error(), instead of exit().
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/grains_hip.h"

namespace grains {

using Transform12 = std::array<double, 12>;  // Transform3 buffer: basis (row major) + origin

struct Kinematics {  // Kinematics<T> (Grains/Component/include/Kinematics.hh)
  std::array<double, 3> translational{}, angular{};
};
struct Torce {  // Torce<T> (Grains/Component/include/Torce.hh)
  std::array<double, 3> torque{}, force{};
};

class ComponentManagerHIP {
 public:
  explicit ComponentManagerHIP(const grains_config_t& cfg) {
    check(grains_create(&cfg, &ctx_), "grains_create");
  }
  ~ComponentManagerHIP() { grains_destroy(ctx_); }
  ComponentManagerHIP(const ComponentManagerHIP&) = delete;
  ComponentManagerHIP& operator=(const ComponentManagerHIP&) = delete;

  // ---- construction (Grains.cpp:91-300) -----------------------------------
  void setShapes(const std::vector<grains_shape_t>& s, const std::vector<double>& vertices = {}) {
    check(grains_set_shapes(ctx_, s.data(), (uint32_t)s.size(), vertices.empty() ? nullptr : vertices.data(),
                            (uint32_t)(vertices.size() / 3)),
          "grains_set_shapes");
  }
  void setContactForceModels(const std::vector<grains_hooke_t>& table, uint32_t n_materials) {
    const uint32_t per = n_materials * (n_materials + 1) / 2;
    check(grains_set_contact_tables(ctx_, table.data(), n_materials, (uint32_t)(table.size() / per)),
          "grains_set_contact_tables");
  }
  void setLinkedCell(const std::vector<grains_sim_t>& sims) {
    check(grains_set_sims(ctx_, sims.data(), (uint32_t)sims.size()), "grains_set_sims");
  }
  // ComponentManager::initializeObstacles (ComponentManagerCPU.cpp:215-231)
  void initializeObstacles(const std::vector<uint32_t>& shape, const std::vector<Transform12>& tr) {
    check(grains_set_obstacles(ctx_, shape.data(), nullptr, tr.empty() ? nullptr : tr[0].data(),
                               (uint32_t)shape.size()),
          "grains_set_obstacles");
  }
  // ComponentManager::initializeParticles (ComponentManagerCPU.cpp:236-252)
  void initializeParticles(const std::vector<uint32_t>& shape, const std::vector<Transform12>& tr) {
    n_ = shape.size();
    check(grains_set_particles(ctx_, shape.data(), nullptr, n_), "grains_set_particles");
    setTransform(tr);
  }

  // ---- state (ComponentManager.hh:40-97) ------------------------------------
  uint32_t getNumberOfParticles() const { return (uint32_t)n_; }
  std::vector<Transform12> getTransform() const {
    std::vector<Transform12> t(n_);
    check(grains_get_transforms(ctx_, t[0].data(), n_), "grains_get_transforms");
    return t;
  }
  void setTransform(const std::vector<Transform12>& t) {
    check(grains_set_transforms(ctx_, t[0].data(), n_), "grains_set_transforms");
  }
  std::vector<Kinematics> getVelocity() const {
    std::vector<double> v(3 * n_), w(3 * n_);
    check(grains_get_state(ctx_, nullptr, nullptr, v.data(), w.data(), n_), "grains_get_state");
    std::vector<Kinematics> k(n_);
    for (uint64_t i = 0; i < n_; ++i)
      for (int c = 0; c < 3; ++c) k[i].translational[c] = v[3 * i + c], k[i].angular[c] = w[3 * i + c];
    return k;
  }
  void setVelocity(const std::vector<Kinematics>& k) {
    std::vector<double> v(3 * n_), w(3 * n_);
    for (uint64_t i = 0; i < n_; ++i)
      for (int c = 0; c < 3; ++c) v[3 * i + c] = k[i].translational[c], w[3 * i + c] = k[i].angular[c];
    check(grains_set_state(ctx_, nullptr, nullptr, v.data(), w.data(), n_), "grains_set_state");
  }
  // torce accumulated by the last step before the integrator reset it
  std::vector<Torce> getTorce() const {
    std::vector<double> f(3 * n_), t(3 * n_);
    check(grains_get_torce(ctx_, f.data(), t.data(), n_), "grains_get_torce");
    std::vector<Torce> r(n_);
    for (uint64_t i = 0; i < n_; ++i)
      for (int c = 0; c < 3; ++c) r[i].force[c] = f[3 * i + c], r[i].torque[c] = t[3 * i + c];
    return r;
  }

  // ---- the time step (GrainsCPU.cpp:55-66) ------------------------------------
  // updateLinks + particle/particle + particle/obstacle contacts; `result`
  // receives the cumulative per-particle contact counts after the step runs
  void detectCollisionAndComputeContactForces(int* result = nullptr) {
    pending_ = true;
    result_ = result;
  }
  // gravity is part of the sim (grains_sim_t.gravity); kept for the call order
  void addExternalForces() {}
  void moveParticles() {
    if (!pending_) throw std::runtime_error("moveParticles before detectCollisionAndComputeContactForces");
    check(grains_step(ctx_, 1), "grains_step");
    pending_ = false;
    if (result_) check(grains_get_contact_count(ctx_, result_, n_), "grains_get_contact_count");
  }
  // n iterations of the loop body at once (no host round trip per step)
  void simulate(uint32_t n_steps, bool graph = true) {
    check(graph ? grains_step_graph(ctx_, n_steps) : grains_step(ctx_, n_steps), "grains_step");
  }

  grains_stats_t stats() const {
    grains_stats_t s;
    check(grains_get_stats(ctx_, &s), "grains_get_stats");
    return s;
  }
  grains_ctx_t* handle() const { return ctx_; }

 private:
  void check(int rc, const char* what) const {
    if (rc != GRAINS_OK)
      throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) +
                               "): " + (ctx_ ? grains_last_error(ctx_) : "no context"));
  }
  grains_ctx_t* ctx_ = nullptr;
  uint64_t n_ = 0;
  bool pending_ = false;
  int* result_ = nullptr;
};

}  // namespace grains
