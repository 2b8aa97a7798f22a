// Drives the reference-shaped C++ classes of phybs.hpp the way the reference application does
// (fvm.cpp:30, 44-47, 175, 187, 216-217) on a 4x4x16 bar, and prints the resulting positions.
//   compat_demo cuda <steps>     CUDAInterface: doDescentOpt(h, 96, ext, nodes) per step
//   compat_demo descent <steps>  DescentOptimizeT over a minimal mesh type: doDescentOpt(96)
//   compat_demo backeuler <steps> CUDAInterface: computeInnerforces, computeGlobalStiffnessMatrix,
//                                 doBackEuler(nodes) per step (fvm.cpp:179-183)
//   compat_demo timeint <steps>  TimeIntegrationT (no GPU): the step of each constructor, then
//                                 <steps> explicit-Euler steps of the twisted bar under a constant
//                                 force, and whether BackEuler throws
//   compat_demo material <TetGen prefix> <out file>
//                                 TetMeshT loads the mesh file by file; IsotropicNeohookeanMaterialT
//                                 evaluates forces, energies, the global and two per-tet stiffness
//                                 matrices and the principal-stretch derivatives at a perturbed
//                                 state; the arrays go to <out file> (int32 tag, int32 count, data)
//   compat_demo descent_stretch <steps>
//                                 DescentOptimizeT over TetMeshT-shaped data with the stretch
//                                 material object: doDescentOpt(96) per step
// Output (other than material): one line per node, "x y z" with 9 significant digits.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <type_traits>
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
  static Mat Zero(int n) { return Zero(n, 1); }
};
typedef phybs::TetMeshT<Mat<float>, Mat<int>, Mat<float>> DemoTetMesh;

// tag 0: float32 array, tag 1: int32 array
template <class T>
void put(FILE *f, const T *p, size_t n) {
  const int tag = std::is_same<T, float>::value ? 0 : 1, cnt = (int)n;
  std::fwrite(&tag, 4, 1, f);
  std::fwrite(&cnt, 4, 1, f);
  std::fwrite(p, sizeof(T), n, f);
}

int material_mode(const char *prefix, const char *out) {
  auto mesh = std::make_shared<DemoTetMesh>();
  mesh->setTetMaterial(1e5f, 0.45f, 1000.0f);
  mesh->initNodesFromFile(std::string(prefix) + ".node");
  mesh->initTetsFromFile(std::string(prefix) + ".ele");
  mesh->initFacesFromFile(std::string(prefix) + ".face");
  const int nn = mesh->getNodesNum(), nt = mesh->getTetsNum();
  // clamp the nodes at the smallest x (the reference's own rule picks none on this mesh)
  float xmin = 1e30f;
  for (int i = 0; i < nn; ++i) xmin = std::min(xmin, mesh->getRestPosition().v[3 * i]);
  mesh->getConstraintIDs().clear();
  for (int i = 0; i < nn; ++i)
    if (mesh->getRestPosition().v[3 * i] == xmin) mesh->getConstraintIDs().push_back(i);
  Mat<float> pos = mesh->getRestPosition();
  for (size_t k = 0; k < pos.v.size(); ++k) pos.v[k] += 0.001f * std::sin(0.7f * (float)k);  // no tet inverts
  mesh->getNodes() = pos;
  mesh->getExternalForces().v[3 * (nn - 1) + 1] = 0.25f;
  phybs::IsotropicNeohookeanMaterialT<DemoTetMesh> mat(mesh);
  Mat<float> f = mat.computeInnerForceFromPos(pos);
  mat.computeElasticEnergyFromPos(pos);
  phybs::CsrMatrix K = mat.computeGlobalStiffnessMatrixFromPos(pos);
  Mat<float> k0 = mat.computeStiffnessMatrixFromPos(0, pos), k1 = mat.computeStiffnessMatrixFromPos(nt - 1, pos);
  const float fh[3] = {1.1f, 0.95f, 0.9f};
  float grad[3], hess[9];
  mat.computeEnergy2FhatGradient(0, fh, grad);
  mat.computeEnergy2FhatHessian(0, fh, hess);
  FILE *o = std::fopen(out, "wb");
  if (!o) return 1;
  put(o, mesh->getRestPosition().data(), 3 * (size_t)nn);
  put(o, mesh->getTets().data(), 4 * (size_t)nt);
  put(o, mesh->getFaces().data(), 3 * (size_t)mesh->getFacesNum());
  put(o, mesh->getMasses().data(), nn);
  put(o, mesh->getInvMasses().data(), nn);
  put(o, mesh->getTetVolumes().data(), nt);
  put(o, mesh->getDmInvs().data(), 9 * (size_t)nt);
  put(o, mesh->getGravities().data(), 3 * (size_t)nn);
  put(o, pos.data(), 3 * (size_t)nn);
  put(o, f.data(), 3 * (size_t)nn);
  put(o, mat.getElasticEnergys().data(), nt);
  put(o, K.rowptr.data(), K.rowptr.size());
  put(o, K.col.data(), K.col.size());
  put(o, K.val.data(), K.val.size());
  put(o, mat.getDiagonalIdx().data(), mat.getDiagonalIdx().size());
  put(o, mat.getKIDinCSRval().data(), mat.getKIDinCSRval().size());
  put(o, k0.data(), 144);
  put(o, k1.data(), 144);
  put(o, grad, 3);
  put(o, hess, 9);
  put(o, mesh->getConstraintIDs().data(), mesh->getConstraintIDs().size());
  std::fclose(o);
  return 0;
}

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
    std::fprintf(stderr, "usage: compat_demo cuda|descent|descent_stretch|backeuler|timeint <steps> | material <prefix> <out>\n");
    return 2;
  }
  if (std::strcmp(argv[1], "material") == 0) {
    if (argc < 4) return 2;
    try {
      return material_mode(argv[2], argv[3]);
    } catch (const phybs::Error &e) {
      std::fprintf(stderr, "%s\n", e.what());
      return 1;
    }
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
  const bool stretch = std::strcmp(argv[1], "descent_stretch") == 0;
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
      if (stretch) {
        // the material object decides the energy: the stretch material's youngs() / poisson() /
        // materialKind() reach the solver
        struct StretchMaterial {
          float youngs() const { return 2e5f; }
          float poisson() const { return 0.4f; }
          int materialKind() const { return PBS_MAT_STRETCH; }
        };
        phybs::DescentOptimizeT<DemoMesh, StretchMaterial> opt(mesh, std::make_shared<StretchMaterial>());
        for (int s = 0; s < steps; ++s) opt.doDescentOpt(96);
      } else {
        phybs::DescentOptimizeT<DemoMesh, DemoMaterial> opt(mesh, std::make_shared<DemoMaterial>());
        for (int s = 0; s < steps; ++s) opt.doDescentOpt(96);
      }
      out = mesh->nodes.v;
    }
  } catch (const phybs::Error &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  for (int i = 0; i < nn; ++i) std::printf("%.9g %.9g %.9g\n", out[3 * i], out[3 * i + 1], out[3 * i + 2]);
  return 0;
}
