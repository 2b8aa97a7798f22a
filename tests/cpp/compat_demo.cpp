// Drives the reference-shaped C++ classes of phybs.hpp the way the reference application does
// (fvm.cpp:30, 44-47, 175, 187, 216-217) on a 4x4x16 bar, and prints the resulting positions.
//   compat_demo cuda <steps>     CUDAInterface: doDescentOpt(h, 96, ext, nodes) per step
//   compat_demo descent <steps>  DescentOptimizeT over a minimal mesh type: doDescentOpt(96)
//   compat_demo backeuler <steps> CUDAInterface: computeInnerforces, computeGlobalStiffnessMatrix,
//                                 doBackEuler(nodes) per step (fvm.cpp:179-183)
//   compat_demo timeint <steps>  TimeIntegrationT (no GPU): the step of each constructor, then
//                                 <steps> explicit-Euler steps of the twisted bar under a constant
//                                 force, and whether BackEuler throws
// Output: one line per node, "x y z" with 9 significant digits.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "phybs.hpp"

namespace {

int steps_arg(const char *s) { return std::atoi(s); }

// Column-major 3xN / 4xN storage with the two members the shim reads.
template <class T>
struct Mat {
  std::vector<T> v;
  T *data() { return v.data(); }
  const T *data() const { return v.data(); }
  void setZero() { std::fill(v.begin(), v.end(), T(0)); }
  static Mat Zero(int rows, int cols) {
    Mat m;
    m.v.assign((size_t)rows * cols, T(0));
    return m;
  }
};

// A mesh type with the TetMesh getters that DescentOptimizeT uses (TetMesh.h:41-56).
struct DemoMesh {
  int nn, nt;
  Mat<float> nodes, rest, vel, ext;
  Mat<int> tets;
  std::vector<int> cons;
  Mat<float> &getNodes() { return nodes; }
  Mat<float> &getRestPosition() { return rest; }
  Mat<float> &getVelocities() { return vel; }
  Mat<float> &getExternalForces() { return ext; }
  Mat<int> &getTets() { return tets; }
  std::vector<int> &getConstraintIDs() { return cons; }
  float getE() { return 1e5f; }
  float getNu() { return 0.45f; }
  int getNodesNum() { return nn; }
  int getTetsNum() { return nt; }
  void resetExternalForce() { ext.setZero(); }
};
struct DemoMaterial {};

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: compat_demo cuda|descent|backeuler|timeint <steps>\n");
    return 2;
  }
  if (std::strcmp(argv[1], "timeint") == 0) {
    const int nx = 4, ny = 4, nz = 16, nn = (nx + 1) * (ny + 1) * (nz + 1), nt = 6 * nx * ny * nz;
    Mat<float> rest, pos, vel, force, mass;
    rest.v.resize(3 * nn);
    std::vector<int> tets(4 * nt);
    phybs::check(pbs_mesh_bar(nx, ny, nz, 0.05f, rest.data(), tets.data()));
    pos.v = rest.v;
    phybs::check(pbs_mesh_twist(nx, ny, nz, 0.05f, 30.0, 0.0, 1, nn, pos.data()));
    vel.v.assign(3 * nn, 0.0f);
    force.v.assign(3 * nn, 0.0f);
    mass.v.assign(nn, 0.0f);
    for (int i = 0; i < nn; ++i) {
      force.v[3 * i + 2] = -9.8f * (1.0f + 0.01f * (float)(i % 7));
      mass.v[i] = 0.125f + 0.001f * (float)(i % 5);
    }
    phybs::TimeIntegrationT<Mat<float>, Mat<float>> a(nn), b(nn, mass), ti(nn, mass, std::vector<int>{0}, rest);
    std::printf("%.9g %.9g %.9g\n", a.getTimeStep(), b.getTimeStep(), ti.getTimeStep());
    ti.setTimeStep(1.0f / 30.0f);
    for (int s = 0; s < steps_arg(argv[2]); ++s) {
      ti.simuExplicit(pos, vel, force, mass);
      pos = ti.getPositions();
      vel = ti.getVelocities();
    }
    int threw = 0;
    try {
      int K = 0;
      ti.BackEuler(pos, vel, force, K);
    } catch (const phybs::Error &e) {
      threw = e.code == PBS_ERR_STATE;
    }
    std::printf("%d\n", threw);
    for (int i = 0; i < nn; ++i) std::printf("%.9g %.9g %.9g\n", pos.v[3 * i], pos.v[3 * i + 1], pos.v[3 * i + 2]);
    return 0;
  }
  const bool cuda = std::strcmp(argv[1], "cuda") == 0;
  const bool backeuler = std::strcmp(argv[1], "backeuler") == 0;
  const int steps = std::atoi(argv[2]);
  const int nx = 4, ny = 4, nz = 16;
  const float edge = 0.05f;
  const int nn = (nx + 1) * (ny + 1) * (nz + 1), nt = 6 * nx * ny * nz;
  std::vector<float> rest(3 * nn);
  std::vector<int> tets(4 * nt);
  phybs::check(pbs_mesh_bar(nx, ny, nz, edge, rest.data(), tets.data()));
  std::vector<float> x0(rest);  // twisted in place
  phybs::check(pbs_mesh_twist(nx, ny, nz, edge, 30.0, 0.0, 1, nn, x0.data()));
  std::vector<int> mask(nn);
  for (int i = 0; i < nn; ++i) mask[i] = rest[3 * i] == 0.0f;
  std::vector<float> out(x0);
  try {
    if (backeuler) {
      phybs::CUDAInterface gpu(nn, x0.data(), rest.data(), mask.data(), nt, tets.data(), 1e5f, 0.45f, 1000.0f, 0,
                               0, nullptr, nullptr, nullptr, nullptr, nullptr);
      gpu.updateNodePositions(x0.data());
      for (int s = 0; s < steps; ++s) {
        gpu.computeInnerforces();
        gpu.computeGlobalStiffnessMatrix();
        gpu.doBackEuler(out.data());
      }
    } else if (cuda) {
      phybs::CUDAInterface gpu(nn, x0.data(), rest.data(), mask.data(), nt, tets.data(), 1e5f, 0.45f, 1000.0f, 0,
                               0, nullptr, nullptr, nullptr, nullptr, nullptr);
      gpu.updateNodePositions(x0.data());
      std::vector<float> ext(3 * nn, 0.0f);
      for (int s = 0; s < steps; ++s) gpu.doDescentOpt(1.0f / 30.0f, 96, ext.data(), out.data());
    } else {
      auto mesh = std::make_shared<DemoMesh>();
      mesh->nn = nn;
      mesh->nt = nt;
      mesh->nodes.v = x0;
      mesh->rest.v = rest;
      mesh->vel.v.assign(3 * nn, 0.0f);
      mesh->ext.v.assign(3 * nn, 0.0f);
      mesh->tets.v = tets;
      for (int i = 0; i < nn; ++i)
        if (mask[i]) mesh->cons.push_back(i);
      phybs::DescentOptimizeT<DemoMesh, DemoMaterial> opt(mesh, std::make_shared<DemoMaterial>());
      for (int s = 0; s < steps; ++s) opt.doDescentOpt(96);
      out = mesh->nodes.v;
    }
  } catch (const phybs::Error &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  for (int i = 0; i < nn; ++i) std::printf("%.9g %.9g %.9g\n", out[3 * i], out[3 * i + 1], out[3 * i + 2]);
  return 0;
}
