// phybs.hpp -- header-only C++ layer over the C ABI (phybs.h) with the reference's class surface.
//
//  * phybs::Solver          RAII owner of a pbs_solver handle; errors become phybs::Error.
//  * phybs::CUDAInterface   the public surface of PhyBS's GPU solver class that its application
//                           calls (FVM/CUDAInterface.h:23-35; used at fvm.cpp:44-47, 101, 187, 216):
//                           same constructor arguments, doDescentOpt, reset, updateNodePositions.
//  * phybs::DescentOptimizeT<Mesh, Material>
//                           the CPU descent class (FVM/DescentOptimize.h:11-18; fvm.cpp:30, 175,
//                           217) over any mesh type with the TetMesh getters it reads
//                           (TetMesh.h:41-56): getNodes(), getRestPosition(), getVelocities(),
//                           getExternalForces(), getTets(), getConstraintIDs(), getE(), getNu(),
//                           getNodesNum(), getTetsNum(), resetExternalForce().  It runs the GPU
//                           solver with the CPU path's semantics (PBS_SEM_REF_CPU).
//  * phybs::TimeIntegrationT<Vec, Mat>
//                           the time-step holder the application reads the step from
//                           (FVM/TimeIntegration.h:13-31; fvm.cpp:19-20, 55, 165, 187).
//
// Define PHYBS_REFERENCE_NAMES before including this header (after the reference's TetMesh.h and
// IsotropicNeohookeanMaterial.h) to get the reference's global names CUDAInterface,
// DescentOptimize and (with Eigen included) TimeIntegration, so that the application's call
// sites compile unchanged.
#ifndef PHYBS_HPP
#define PHYBS_HPP

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "phybs.h"

namespace phybs {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
  if (rc != PBS_OK) throw Error(rc, std::string("phybs: ") + pbs_last_error());
}

inline pbs_params default_params() {
  pbs_params p;
  check(pbs_default_params(&p));
  return p;
}

// RAII handle.  Host arrays are AoS xyz floats (3 per node) and 4 int32 per tet.
class Solver {
 public:
  Solver(int num_nodes, const float *nodes, const float *rest, const int *mask, int num_tets, const int *tets,
         const pbs_params &p) {
    check(pbs_create(num_nodes, nodes, rest, mask, num_tets, tets, &p, &h_));
    nn_ = num_nodes;
    nt_ = num_tets;
  }
  ~Solver() { pbs_destroy(h_); }
  Solver(const Solver &) = delete;
  Solver &operator=(const Solver &) = delete;

  pbs_solver *handle() const { return h_; }
  int num_nodes() const { return nn_; }
  int num_tets() const { return nt_; }
  pbs_params params() const {
    pbs_params p;
    check(pbs_get_params(h_, &p));
    return p;
  }
  void set_params(const pbs_params &p) { check(pbs_set_params(h_, &p)); }
  void reset() { check(pbs_reset(h_)); }
  void set_positions(const float *xyz) { check(pbs_set_positions(h_, xyz)); }
  void positions(float *xyz) const { check(pbs_get_positions(h_, xyz)); }
  void set_velocities(const float *v, const float *v_last) { check(pbs_set_velocities(h_, v, v_last)); }
  void velocities(float *v, float *v_last) const { check(pbs_get_velocities(h_, v, v_last)); }
  pbs_step_stats step(const float *ext = nullptr, float *out = nullptr) {
    pbs_step_stats st;
    check(pbs_step(h_, ext, out, &st));
    return st;
  }
  void forces(const float *pos, float *tet_forces, float *node_forces, float *tet_energy) {
    check(pbs_compute_forces(h_, pos, tet_forces, node_forces, tet_energy));
  }
  void hessian_diag(const float *pos, float *node_hd, float *tet_hd) {
    check(pbs_compute_hessian_diag(h_, pos, node_hd, tet_hd));
  }

 private:
  pbs_solver *h_ = nullptr;
  int nn_ = 0, nt_ = 0;
};

// The GPU solver class of the reference application.  The CSR arguments describe the global
// stiffness matrix of the reference's implicit-Euler path; the descent step does not use them
// and they are ignored.  Gravity is (0, -9.8, 0) and dt 0.03, as in the reference.
class CUDAInterface {
 public:
  CUDAInterface(int num_nodes, const float *nodes, const float *restPoses, const int *constraintsMask, int num_tets,
                const int *tets, const float youngs, const float nu, const float density, int csr_nnz, int csr_m,
                float *csr_val, int *csr_row, int *csr_col, int *csr_diagonalIdx, int *csr_kIDinCSRval)
      : nn_(num_nodes) {
    (void)csr_nnz, (void)csr_m, (void)csr_val, (void)csr_row, (void)csr_col, (void)csr_diagonalIdx,
        (void)csr_kIDinCSRval;
    pbs_params p = default_params();
    p.youngs = youngs;
    p.nu = nu;
    p.density = density;
    solver_.reset(new Solver(num_nodes, nodes, restPoses, constraintsMask, num_tets, tets, p));
  }
  // h is ignored, as in the reference (CUDAInterface.cu:911 steps with its member m_timestep).
  void doDescentOpt(float h, int iterations, float *hostExternForces, float *hostNode) {
    check(pbs_do_descent_opt(solver_->handle(), h, iterations, hostExternForces, hostNode));
  }
  void reset() { solver_->reset(); }
  void updateNodePositions(const float *hostNode) { solver_->set_positions(hostNode); }
  // The implicit path as fvm.cpp:179-183 drives it: computeInnerforces(), then
  // computeGlobalStiffnessMatrix(), then doBackEuler(nodes).  pbs_implicit_step evaluates the
  // forces and the stiffness matrix itself, so the first two calls only mark the step as started.
  void computeInnerforces() { forces_ready_ = true; }
  void computeGlobalStiffnessMatrix() { stiffness_ready_ = true; }
  void doBackEuler(float *hostNode) {
    if (!forces_ready_ || !stiffness_ready_)
      throw Error(PBS_ERR_STATE, "doBackEuler: call computeInnerforces and computeGlobalStiffnessMatrix first");
    pbs_implicit_params ip;
    check(pbs_implicit_default_params(&ip));
    check(pbs_implicit_step(solver_->handle(), &ip, nullptr, hostNode, nullptr));
    forces_ready_ = stiffness_ready_ = false;
  }
  Solver &solver() { return *solver_; }

 private:
  int nn_;
  std::unique_ptr<Solver> solver_;
  bool forces_ready_ = false, stiffness_ready_ = false;
};

// The CPU descent class of the reference, on the GPU.  Each doDescentOpt(iterations) takes the
// mesh's current nodes, velocities and external forces, runs one implicit-Euler step with the
// CPU path's semantics, writes the new nodes and velocities back into the mesh and clears its
// external forces (DescentOptimize.cpp:91-188).
template <class Mesh, class Material>
class DescentOptimizeT {
 public:
  DescentOptimizeT(std::shared_ptr<Mesh> tetMesh, std::shared_ptr<Material> isoMaterial, float density = 1000.0f)
      : mesh_(tetMesh), material_(isoMaterial) {
    const int nn = mesh_->getNodesNum(), nt = mesh_->getTetsNum();
    std::vector<int> mask(nn, 0);
    for (int id : mesh_->getConstraintIDs()) mask[id] = 1;
    pbs_params p = default_params();
    p.semantics = PBS_SEM_REF_CPU;
    p.youngs = mesh_->getE();
    p.nu = mesh_->getNu();
    p.density = density;
    solver_.reset(new Solver(nn, mesh_->getNodes().data(), mesh_->getRestPosition().data(), mask.data(), nt,
                             mesh_->getTets().data(), p));
    v_last_.assign(3 * (size_t)nn, 0.0f);
    reset();
  }
  // DescentOptimize::reset (DescentOptimize.cpp:30-34): the current mesh velocities become the
  // previous-step velocities.
  void reset() {
    const float *v = mesh_->getVelocities().data();
    v_last_.assign(v, v + 3 * (size_t)mesh_->getNodesNum());
  }
  void doDescentOpt(int iterations) {
    pbs_params p = solver_->params();
    if (p.iterations != iterations) {
      p.iterations = iterations;
      solver_->set_params(p);
    }
    solver_->set_positions(mesh_->getNodes().data());
    solver_->set_velocities(mesh_->getVelocities().data(), v_last_.data());
    const float *v = mesh_->getVelocities().data();
    std::vector<float> v_prev(v, v + 3 * (size_t)mesh_->getNodesNum());
    solver_->step(mesh_->getExternalForces().data(), mesh_->getNodes().data());
    solver_->velocities(mesh_->getVelocities().data(), nullptr);
    v_last_.swap(v_prev);
    mesh_->resetExternalForce();
  }
  Solver &solver() { return *solver_; }

 private:
  std::shared_ptr<Mesh> mesh_;
  std::shared_ptr<Material> material_;
  std::unique_ptr<Solver> solver_;
  std::vector<float> v_last_;
};

// The reference's time-integration object as the application uses it: the time step, with the
// defaults of the reference's three constructors (1e-6, 2e-3 and 0.03 s; TimeIntegration.cpp:6-30),
// explicit Euler (simuExplicit, TimeIntegration.cpp:57-67) and the position / velocity getters.
// Vec and Mat are the caller's vector and 3 x N column-major matrix types (Eigen::VectorXf,
// Eigen::MatrixXf): data() and Mat::Zero(rows, cols) are all this class uses.  BackEuler with a
// stiffness matrix the caller assembled on the CPU (TimeIntegration.cpp:69-118, factorised with
// MKL Pardiso) is not provided and throws PBS_ERR_STATE: the GPU backward-Euler step is
// CUDAInterface::computeInnerforces / computeGlobalStiffnessMatrix / doBackEuler.
template <class Vec, class Mat>
class TimeIntegrationT {
 public:
  explicit TimeIntegrationT(int n_nodes) : t_(1e-6f), n_(n_nodes) { zero(); }
  TimeIntegrationT(int n_nodes, const Vec &) : t_(2e-3f), n_(n_nodes) { zero(); }
  TimeIntegrationT(int n_nodes, const Vec &, const std::vector<int> &, const Mat &) : t_(0.03f), n_(n_nodes) {
    zero();
  }
  // v' = v + t f / m, x' = x + t v' per node (pos, vel, force: 3 x N; mass: N)
  void simuExplicit(const Mat &pos, const Mat &vel, const Mat &force, const Vec &mass) {
    const float *p = pos.data(), *v = vel.data(), *f = force.data(), *m = mass.data();
    float *xp = positions_.data(), *vp = velocities_.data();
    for (int i = 0; i < n_; ++i)
      for (int d = 0; d < 3; ++d) {
        const size_t k = 3 * (size_t)i + d;
        vp[k] = v[k] + t_ * f[k] / m[i];
        xp[k] = p[k] + vp[k] * t_;
      }
  }
  template <class Sparse>
  void BackEuler(Mat &, Mat &, Mat &, Sparse &) {
    throw Error(PBS_ERR_STATE,
                "phybs: TimeIntegration::BackEuler with a CPU-assembled K is not provided; "
                "use CUDAInterface::computeInnerforces/computeGlobalStiffnessMatrix/doBackEuler");
  }
  void setTimeStep(float t) { t_ = t; }
  Mat getPositions() { return positions_; }
  Mat getVelocities() { return velocities_; }
  float &getTimeStep() { return t_; }

 private:
  void zero() {
    positions_ = Mat::Zero(3, n_);
    velocities_ = Mat::Zero(3, n_);
  }
  float t_;
  int n_;
  Mat positions_, velocities_;
};

}  // namespace phybs

#ifdef PHYBS_REFERENCE_NAMES
using phybs::CUDAInterface;
typedef phybs::DescentOptimizeT<TetMesh, IsotropicNeohookeanMaterial> DescentOptimize;
#ifdef EIGEN_WORLD_VERSION
typedef phybs::TimeIntegrationT<Eigen::VectorXf, Eigen::MatrixXf> TimeIntegration;
#endif
#endif

#endif  // PHYBS_HPP
