// Drives the C++ ComponentManagerHIP adapter the way the reference's
// GrainsCPU::simulate loop drives its ComponentManager
// (Grains/Base/src/GrainsCPU.cpp:46-70): two spheres dropped on a floor.
// Prints "ok <contacts> <z0> <z1>" on success.
#include <cmath>
#include <cstdio>

#include "../../grains3dgpu_amd/host/ComponentManagerHIP.hh"

int main() {
  grains_config_t cfg{};
  cfg.dt = 1e-5;
  cfg.precision = GRAINS_FP64;
  grains::ComponentManagerHIP cm(cfg);
  grains_shape_t sphere{}, floor{};
  sphere.kind = GRAINS_SPHERE;
  sphere.dims[0] = 0.5e-3;
  sphere.crust = 2.5e-5;
  sphere.density = 2500.0;
  floor.kind = GRAINS_RECTANGLE;
  floor.dims[0] = floor.dims[1] = 5e-3;
  cm.setShapes({sphere, floor});
  grains_hooke_t h{10.0, 0.3, 0.0, 0.3, 0.0};
  cm.setContactForceModels({h}, 1);
  grains_sim_t sim{};
  for (int k = 0; k < 3; ++k) sim.domain_min[k] = -5e-3, sim.domain_max[k] = 5e-3;
  sim.gravity[2] = -9.81;
  cm.setLinkedCell({sim});
  grains::Transform12 I{1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
  grains::Transform12 fl = I;
  fl[11] = -1e-3;
  cm.initializeObstacles({1}, {fl});
  grains::Transform12 a = I, b = I;
  a[9] = -1.5e-3, a[11] = -1e-3 + 0.5e-3;
  b[9] = 1.5e-3, b[11] = -1e-3 + 0.6e-3;
  cm.initializeParticles({0, 0}, {a, b});
  int result[2] = {0, 0};
  for (int step = 0; step < 20000; ++step) {
    cm.detectCollisionAndComputeContactForces(result);
    cm.addExternalForces();
    cm.moveParticles();
  }
  auto tr = cm.getTransform();
  const double m = 2500.0 * 4.0 / 3.0 * M_PI * 0.5e-3 * 0.5e-3 * 0.5e-3;
  const double rest = -1e-3 + 0.5e-3 - m * 9.81 / 10.0;
  const double z0 = tr[0][11], z1 = tr[1][11];
  const bool ok = std::fabs(z0 - rest) < 1e-9 && std::fabs(z1 - rest) < 1e-9 && result[0] == 0;
  std::printf("%s %d %.12e %.12e\n", ok ? "ok" : "FAIL", result[0] + result[1], z0, z1);
  return ok ? 0 : 1;
}
