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
//  * phybs::TetMeshT<MatXf, MatXi, VecXf>
//                           a TetMesh-shaped mesh (FVM/TetMesh.h:15-61): initNodesFromFile /
//                           initTetsFromFile / initFacesFromFile, setTetMaterial, reset, the
//                           reference-returning getters; rest data (volumes, Dm^-1, masses,
//                           gravity) from the GPU setup.
//  * phybs::IsotropicNeohookeanMaterialT<Mesh> (and IsotropicStretchMaterialT<Mesh>)
//                           the material calls of FVM/IsotropicNeohookeanMaterial.h:19-20 and
//                           FVM/IsotropicMaterial.h:30-57 on the GPU: computeInnerForceFromPos,
//                           computeElasticEnergyFromPos / getElasticEnergys,
//                           computeGlobalStiffnessMatrixFromPos, computeStiffnessMatrixFromPos,
//                           getDiagonalIdx, getKIDinCSRval, computeEnergy2FhatGradient / Hessian.
//
// MatXf / MatXi / VecXf are the caller's dense types (Eigen::MatrixXf, MatrixXi, VectorXf):
// column-major storage behind data(), and static Zero(rows, cols) / Zero(n) are all these classes
// use.  The global stiffness matrix comes back as phybs::CsrMatrix (the reference's RowMajor CSR
// arrays), or as Eigen::SparseMatrix<float, Eigen::RowMajor> when Eigen is included first.
//
// Define PHYBS_REFERENCE_NAMES before including this header (after the reference's TetMesh.h and
// IsotropicNeohookeanMaterial.h) to get the reference's global names CUDAInterface,
// DescentOptimize and (with Eigen included) TimeIntegration, so that the application's call
// sites compile unchanged.
#ifndef PHYBS_HPP
#define PHYBS_HPP

#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
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
  void save_state(const char *path) { check(pbs_state_save(h_, path)); }
  void load_state(const char *path) { check(pbs_state_load(h_, path)); }
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

namespace detail {
// QString (toStdString) or anything a std::string can be made from
template <class S>
auto to_path(const S &s, int) -> decltype(s.toStdString()) {
  return s.toStdString();
}
template <class S>
std::string to_path(const S &s, long) {
  return std::string(s);
}
template <class S>
std::string path(const S &s) {
  return to_path(s, 0);
}
// optional members of a caller's mesh / material type
template <class M>
auto density_of(M &m, int) -> decltype((float)m.getDensity()) {
  return m.getDensity();
}
template <class M>
float density_of(M &, long) {
  return 1000.0f;
}
template <class T>
auto material_kind(T &m, int) -> decltype((int)m.materialKind()) {
  return m.materialKind();
}
template <class T>
int material_kind(T &, long) {
  return PBS_MAT_NEOHOOKEAN;
}
template <class T, class Mesh>
auto youngs_of(T &m, Mesh &, int) -> decltype((float)m.youngs()) {
  return m.youngs();
}
template <class T, class Mesh>
float youngs_of(T &, Mesh &mesh, long) {
  return mesh.getE();
}
template <class T, class Mesh>
auto poisson_of(T &m, Mesh &, int) -> decltype((float)m.poisson()) {
  return m.poisson();
}
template <class T, class Mesh>
float poisson_of(T &, Mesh &mesh, long) {
  return mesh.getNu();
}
}  // namespace detail

// The reference's RowMajor sparse matrix as plain CSR arrays (Eigen's outer / inner / value
// arrays of a compressed SparseMatrix<float, RowMajor>).
struct CsrMatrix {
  int rows = 0, cols = 0;
  std::vector<int> rowptr, col;
  std::vector<float> val;
};
template <class Sparse>
struct SparseFromCsr;  // Sparse make(const CsrMatrix &)
template <>
struct SparseFromCsr<CsrMatrix> {
  static CsrMatrix make(const CsrMatrix &c) { return c; }
};
#ifdef EIGEN_WORLD_VERSION
template <>
struct SparseFromCsr<Eigen::SparseMatrix<float, Eigen::RowMajor>> {
  static Eigen::SparseMatrix<float, Eigen::RowMajor> make(const CsrMatrix &c) {
    return Eigen::Map<const Eigen::SparseMatrix<float, Eigen::RowMajor>>(c.rows, c.cols, (int)c.val.size(),
                                                                          c.rowptr.data(), c.col.data(),
                                                                          c.val.data());
  }
};
#endif

// A TetMesh-shaped mesh (TetMesh.h:15-61).  The reference's constructor loads hard-coded model
// paths (TetMesh.cpp:153-219); here the caller loads with initNodesFromFile (positions times 0.1,
// constrained nodes by the reference's rule z > 0 && y > 8.5 after scaling, TetMesh.cpp:36-40),
// initTetsFromFile and initFacesFromFile, and may edit getConstraintIDs() (ascending).  The rest
// data -- volumes, Dm^-1, lumped masses (constrained nodes 1e8, inverse mass 0), gravity
// (0, -9.8 m, 0) with the mass before the override -- come from the GPU setup on first use and
// after setTetMaterial / a new file.
template <class MatXf, class MatXi, class VecXf>
class TetMeshT {
 public:
  TetMeshT() = default;
  template <class S>
  void initNodesFromFile(const S &filename) {
    const std::string f = detail::path(filename);
    int n = 0;
    check(pbs_tetgen_read_nodes(f.c_str(), 0.1f, &n, nullptr));
    rest_ = MatXf::Zero(3, n);
    check(pbs_tetgen_read_nodes(f.c_str(), 0.1f, &n, rest_.data()));
    nn_ = n;
    cons_.clear();
    const float *r = rest_.data();
    for (int i = 0; i < n; ++i)
      if (r[3 * i + 2] > 0 && r[3 * i + 1] > 8.5f) cons_.push_back(i);
    nodes_ = rest_;
    vel_ = MatXf::Zero(3, n);
    ext_ = MatXf::Zero(3, n);
    forces_ = MatXf::Zero(3, n);
    stale_ = true;
  }
  template <class S>
  void initTetsFromFile(const S &filename) {
    const std::string f = detail::path(filename);
    int n = 0;
    check(pbs_tetgen_read_tets(f.c_str(), &n, nullptr));
    tets_ = MatXi::Zero(4, n);
    check(pbs_tetgen_read_tets(f.c_str(), &n, tets_.data()));
    nt_ = n;
    stale_ = true;
  }
  template <class S>
  void initFacesFromFile(const S &filename) {
    const std::string f = detail::path(filename);
    int n = 0;
    check(pbs_tetgen_read_faces(f.c_str(), &n, nullptr));
    faces_ = MatXi::Zero(3, n);
    check(pbs_tetgen_read_faces(f.c_str(), &n, faces_.data()));
    nf_ = n;
  }
  // .node import without the 0.1 scale (TetMesh.cpp:70-94) and export (TetMesh.cpp:558-570)
  template <class S>
  void updateNodesFromFile(const S &filename) {
    check(pbs_node_read(detail::path(filename).c_str(), nn_, nodes_.data()));
  }
  template <class S>
  void writeNodesToFile(const S &filename) {
    check(pbs_node_write(detail::path(filename).c_str(), nn_, nodes_.data()));
  }
  // TetMesh.cpp:367-376 (the density enters the lumped masses)
  void setTetMaterial(float e, float nu, float den) {
    E_ = e;
    nu_ = nu;
    density_ = den;
    stale_ = true;
  }
  // TetMesh.cpp:221-227
  void reset() {
    nodes_ = rest_;
    vel_ = MatXf::Zero(3, nn_);
    ext_ = MatXf::Zero(3, nn_);
    forces_ = MatXf::Zero(3, nn_);
  }
  void resetExternalForce() { ext_ = MatXf::Zero(3, nn_); }

  MatXf &getNodes() { return nodes_; }
  MatXf &getVelocities() { return vel_; }
  MatXf &getForces() { return forces_; }
  MatXf &getExternalForces() { return ext_; }
  MatXf &getRestPosition() { return rest_; }
  MatXi &getTets() { return tets_; }
  MatXi &getFaces() { return faces_; }
  std::vector<int> &getConstraintIDs() { return cons_; }
  int getNodesNum() { return nn_; }
  int getTetsNum() { return nt_; }
  int getFacesNum() { return nf_; }
  float getE() { return E_; }
  float getNu() { return nu_; }
  float getDensity() { return density_; }
  int getNodeGlobalIDinTet(int tetID, int localID) { return tets_.data()[4 * (size_t)tetID + localID]; }
  VecXf &getMasses() { return rest_data(), mass_; }
  VecXf &getInvMasses() { return rest_data(), inv_mass_; }
  MatXf &getGravities() { return rest_data(), grav_; }
  VecXf &getTetVolumes() { return rest_data(), vol_; }
  float getTetVolume(int i) { return rest_data(), vol_.data()[i]; }
  MatXf &getDmInvs() { return rest_data(), dminv_; }  // 3 x 3T, block t = Dm^-1 of tet t

  // interactive input (TetMesh.cpp:393-418, 446-470): the face hit by a ray, and the external
  // forces of dragging its ring
  int pickFacebyRay(const float *orig, const float *direct) {
    int face = -1;
    check(pbs_pick_face(nn_, nodes_.data(), nf_, faces_.data(), orig, direct, &face));
    return face;
  }
  void dragFaceRing(int faceID, const float *dragline) {
    if (ring_off_.empty()) {
      const int64_t len = pbs_face_ring(nn_, nf_, faces_.data(), nullptr, nullptr);
      if (len < 0) check((int)len);
      ring_off_.assign(nf_ + 1, 0);
      ring_.assign((size_t)len, 0);
      check((int)std::min<int64_t>(0, pbs_face_ring(nn_, nf_, faces_.data(), ring_off_.data(), ring_.data())));
    }
    rest_data();
    check(pbs_drag_face_ring(nn_, nodes_.data(), mass_.data(), nf_, faces_.data(), ring_off_.data(), ring_.data(),
                             faceID, dragline, ext_.data()));
  }

 private:
  void rest_data() {
    if (!stale_) return;
    std::vector<int> mask(nn_, 0);
    for (int id : cons_) mask[id] = 1;
    pbs_params p = default_params();
    p.youngs = E_;
    p.nu = nu_;
    p.density = density_;
    Solver s(nn_, nullptr, rest_.data(), mask.data(), nt_, tets_.data(), p);
    std::vector<float> dm(9 * (size_t)nt_);
    vol_ = VecXf::Zero(nt_);
    mass_ = VecXf::Zero(nn_);
    grav_ = MatXf::Zero(3, nn_);
    check(pbs_get_rest_data(s.handle(), dm.data(), vol_.data(), mass_.data(), grav_.data()));
    dminv_ = MatXf::Zero(3, 3 * nt_);  // column-major 3 x 3T: (r, 3t + c) = Dm^-1_t[r][c]
    float *d = dminv_.data();
    for (int t = 0; t < nt_; ++t)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) d[(size_t)(3 * t + c) * 3 + r] = dm[9 * (size_t)t + 3 * r + c];
    inv_mass_ = VecXf::Zero(nn_);
    for (int i = 0; i < nn_; ++i) inv_mass_.data()[i] = mask[i] ? 0.0f : 1.0f / mass_.data()[i];
    stale_ = false;
  }
  int nn_ = 0, nt_ = 0, nf_ = 0;
  float E_ = 1e5f, nu_ = 0.45f, density_ = 1000.0f;
  MatXf nodes_, rest_, vel_, ext_, forces_, grav_, dminv_;
  MatXi tets_, faces_;
  VecXf mass_, inv_mass_, vol_;
  std::vector<int> cons_;
  std::vector<int64_t> ring_off_;
  std::vector<int> ring_;
  bool stale_ = true;
};

// The material classes' calls (IsotropicNeohookeanMaterial.h:19-20, IsotropicMaterial.h:30-57) on
// the GPU for a mesh with TetMesh's getters.  Kind selects the energy: PBS_MAT_NEOHOOKEAN
// (IsotropicNeohookeanMaterial) or PBS_MAT_STRETCH (psi = sum mu (s_i - 1)^2 + lambda/2 (J - 1)^2,
// no reference class).  E, nu and the density are the mesh's at construction, as the reference
// material reads them in its constructor (IsotropicNeohookeanMaterial.cpp:15-20); Sparse is the
// type computeGlobalStiffnessMatrixFromPos returns.
template <class Mesh, int Kind = PBS_MAT_NEOHOOKEAN, class Sparse = CsrMatrix>
class IsotropicMaterialT {
 public:
  typedef typename std::decay<decltype(std::declval<Mesh &>().getNodes())>::type MatXf;
  typedef typename std::decay<decltype(std::declval<Mesh &>().getMasses())>::type VecXf;

  explicit IsotropicMaterialT(std::shared_ptr<Mesh> tetMesh) : mesh_(tetMesh) {
    Mesh &m = *mesh_;
    E_ = m.getE();
    nu_ = m.getNu();
    density_ = detail::density_of(m, 0);
    nn_ = m.getNodesNum();
    nt_ = m.getTetsNum();
    std::vector<int> mask(nn_, 0);
    for (int id : m.getConstraintIDs()) mask[id] = 1;
    pbs_params p = default_params();
    p.youngs = E_;
    p.nu = nu_;
    p.density = density_;
    p.material = Kind;
    p.semantics = PBS_SEM_REF_CPU;  // the CPU material's energy clamp E > 1e-6 ? E : 0
    solver_.reset(new Solver(nn_, nullptr, m.getRestPosition().data(), mask.data(), nt_, m.getTets().data(), p));
    // the stiffness pattern, getDiagonalIdx and getKIDinCSRval (allocateGlobalStiffnessMatrix,
    // IsotropicMaterial.cpp:554-644)
    const int64_t nnz = pbs_csr_pattern(nn_, nt_, m.getTets().data(), nullptr, nullptr, nullptr, nullptr);
    if (nnz < 0) check((int)nnz);
    K_.rows = K_.cols = 3 * nn_;
    K_.rowptr.assign(3 * (size_t)nn_ + 1, 0);
    K_.col.assign((size_t)nnz, 0);
    K_.val.assign((size_t)nnz, 0.0f);
    diag_.assign(3 * (size_t)nn_, 0);
    kid_.assign(48 * (size_t)nt_, 0);
    check((int)std::min<int64_t>(0, pbs_csr_pattern(nn_, nt_, m.getTets().data(), K_.rowptr.data(), K_.col.data(),
                                                    diag_.data(), kid_.data())));
    energies_ = VecXf::Zero(nt_);
  }
  float youngs() const { return E_; }
  float poisson() const { return nu_; }
  float density() const { return density_; }
  int materialKind() const { return Kind; }

  // Forces the descent reads (IsotropicNeohookeanMaterial.cpp:93-114): gravity + external forces
  // + the elastic forces at pos on every node; constrained nodes instead carry the spring
  // 1e8 (rest - current mesh nodes) (TetMesh.cpp:580-586).  Also stored in the mesh's getForces().
  MatXf computeInnerForceFromPos(const MatXf &pos) {
    Mesh &m = *mesh_;
    MatXf f = MatXf::Zero(3, nn_);
    check(pbs_compute_forces(solver_->handle(), pos.data(), nullptr, f.data(), nullptr));
    std::vector<float> g(3 * (size_t)nn_);
    check(pbs_get_rest_data(solver_->handle(), nullptr, nullptr, nullptr, g.data()));
    float *o = f.data();
    const float *ext = m.getExternalForces().data();
    for (size_t i = 0; i < 3 * (size_t)nn_; ++i) o[i] = (g[i] + ext[i]) + o[i];
    const float *rest = m.getRestPosition().data(), *cur = m.getNodes().data();
    for (int id : m.getConstraintIDs())
      for (int d = 0; d < 3; ++d) o[3 * (size_t)id + d] = 1e8f * (rest[3 * (size_t)id + d] - cur[3 * (size_t)id + d]);
    m.getForces() = f;
    return f;
  }
  // IsotropicNeohookeanMaterial.cpp:116-131: psi * vol per tet, E > 1e-6 ? E : 0
  void computeElasticEnergyFromPos(const MatXf &pos) {
    check(pbs_compute_forces(solver_->handle(), pos.data(), nullptr, nullptr, energies_.data()));
  }
  VecXf &getElasticEnergys() { return energies_; }
  // K = -df/dx of the elastic forces at pos, assembled over all tets (IsotropicMaterial.cpp:646-682)
  Sparse computeGlobalStiffnessMatrixFromPos(const MatXf &pos) {
    solver_->set_positions(pos.data());
    std::vector<double> v(K_.val.size());
    check(pbs_stiffness_matrix(solver_->handle(), (int64_t)v.size(), v.data()));
    for (size_t i = 0; i < v.size(); ++i) K_.val[i] = (float)v[i];
    return SparseFromCsr<Sparse>::make(K_);
  }
  // 12 x 12 K of one tet (IsotropicMaterial.cpp:503-552); rows / columns 3 * corner + axis
  MatXf computeStiffnessMatrixFromPos(int tetID, const MatXf &pos) {
    float k[144];
    check(pbs_tet_stiffness(solver_->handle(), pos.data(), 1, &tetID, k));
    MatXf K = MatXf::Zero(12, 12);
    for (int r = 0; r < 12; ++r)
      for (int c = 0; c < 12; ++c) K.data()[12 * c + r] = k[12 * r + c];  // column-major
    return K;
  }
  std::vector<int> &getDiagonalIdx() { return diag_; }
  std::vector<int> &getKIDinCSRval() { return kid_; }
  const CsrMatrix &getGlobalK() { return K_; }  // the last matrix computeGlobalStiffnessMatrixFromPos formed

  // Energy derivatives in the principal stretches (IsotropicNeohookeanMaterial.cpp:50-91): for
  // psi = mu/2 (sum s_i^2 - 3) - mu ln J + lambda/2 (ln J)^2, J = s0 s1 s2,
  //   d psi / d s_i = mu s_i + (lambda ln J - mu) / s_i,
  //   d2 psi / d s_i^2 = mu + (mu + lambda - lambda ln J) / s_i^2, d2 psi / d s_i d s_j = lambda / (s_i s_j);
  // the stretch energy's: 2 mu (s_i - 1) + lambda (J - 1) J / s_i, 2 mu + lambda (J / s_i)^2,
  // lambda (2 J - 1) J / (s_i s_j).  hessian: row-major 3 x 3.
  void computeEnergy2FhatGradient(int, const float *Fhats, float *gradient) {
    const float J = Fhats[0] * Fhats[1] * Fhats[2];
    for (int i = 0; i < 3; ++i)
      gradient[i] = Kind == PBS_MAT_NEOHOOKEAN ? mu() * Fhats[i] + (lam() * std::log(J) - mu()) / Fhats[i]
                                               : 2 * mu() * (Fhats[i] - 1) + lam() * (J - 1) * J / Fhats[i];
  }
  void computeEnergy2FhatHessian(int, const float *Fhats, float *hessian) {
    const float J = Fhats[0] * Fhats[1] * Fhats[2], lj = std::log(J);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        float h;
        if (Kind == PBS_MAT_NEOHOOKEAN)
          h = i == j ? mu() + (mu() + lam() - lam() * lj) / (Fhats[i] * Fhats[i]) : lam() / (Fhats[i] * Fhats[j]);
        else
          h = i == j ? 2 * mu() + lam() * (J / Fhats[i]) * (J / Fhats[i]) : lam() * (2 * J - 1) * J / (Fhats[i] * Fhats[j]);
        hessian[3 * i + j] = h;
      }
  }
  Solver &solver() { return *solver_; }

 private:
  // mu = E / (2 (1 + nu)), lambda = 2 mu / (1 - 2 nu) (IsotropicNeohookeanMaterial.cpp:15-20)
  float mu() const { return (float)(0.5 * (double)E_ / (double)(1 + nu_)); }
  float lam() const { return mu() * 2 / (1 - 2 * nu_); }
  std::shared_ptr<Mesh> mesh_;
  std::unique_ptr<Solver> solver_;
  float E_, nu_, density_;
  int nn_, nt_;
  CsrMatrix K_;
  std::vector<int> diag_, kid_;
  VecXf energies_;
};
template <class Mesh, class Sparse = CsrMatrix>
using IsotropicNeohookeanMaterialT = IsotropicMaterialT<Mesh, PBS_MAT_NEOHOOKEAN, Sparse>;
template <class Mesh, class Sparse = CsrMatrix>
using IsotropicStretchMaterialT = IsotropicMaterialT<Mesh, PBS_MAT_STRETCH, Sparse>;

// The CPU descent class of the reference, on the GPU.  Each doDescentOpt(iterations) takes the
// mesh's current nodes, velocities and external forces, runs one implicit-Euler step with the
// CPU path's semantics, writes the new nodes and velocities back into the mesh and clears its
// external forces (DescentOptimize.cpp:91-188).  The energy, E, nu and density are the
// material's when it provides materialKind() / youngs() / poisson() / density() (the
// IsotropicMaterialT classes above do), else neo-Hookean with the mesh's getE() / getNu() and
// the density argument (the reference's IsotropicNeohookeanMaterial reads E and nu from the mesh).
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
    p.material = detail::material_kind(*material_, 0);
    p.youngs = detail::youngs_of(*material_, *mesh_, 0);
    p.nu = detail::poisson_of(*material_, *mesh_, 0);
    p.density = material_density(*material_, density, 0);
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
  template <class T>
  static auto material_density(T &m, float, int) -> decltype((float)m.density()) {
    return m.density();
  }
  template <class T>
  static float material_density(T &, float d, long) {
    return d;
  }
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
