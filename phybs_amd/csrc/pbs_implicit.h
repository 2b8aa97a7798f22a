// pbs_implicit.h -- backward-Euler step (the reference's GPU implicit path): per-node assembly
// of the global stiffness matrix and right-hand side, and a Jacobi-preconditioned conjugate-
// gradient solve, all on the device.  Kernels in pbs_implicit.hip, host side in pbs_implicit.cpp.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "pbs_internal.h"

namespace pbs {

// Matrix layout: 3x3 blocks, one block row per node (solver order), stored as SELL-64: rows are
// grouped in slices of 64 (one wavefront), a slice is as wide as its longest row, and slot j of
// the slice's row l sits at (slice_off[s] + j) * 64 + l; block values are spread over 9
// consecutive slot groups so that each of the 9 loads of a wavefront is one contiguous 256-byte
// (FP32) access.  Padding slots have column = the row itself and zero values.
struct IeCtrl {
  double rz, pAp, alpha, beta, rr, bb, rr0, rz0;  // bb = b.D^-1.b
  int k, done, converged, indefinite;
};

struct IeDev {
  int n, nslice, maxit, scalar;    // scalar: one value per slot (block = value x identity)
  double tol2;                     // (relative residual tolerance)^2
  const int *slice_off;            // nslice + 1, in slot groups of 64
  const int *row_len;              // N
  const int *bcol;                 // slots
  const int *diag_slot;            // N: slot index of each row's diagonal block
  void *A;                         // 9 x slots values of the solver's real type
  double *b, *x, *r, *z, *p, *Ap, *dinv;  // 3N, node-major (x, y, z of a node together)
  double *part;                    // 4 partial sums per 256-row block (per 256 entries in updates)
  IeCtrl *ctrl;
};

// Constants of one assembly (scales of the system matrix, fixed-node spring, damping).
struct IeAsm {
  float h, alpha, beta;  // time step, mass and stiffness damping coefficients
  int raw;               // 1: store K itself (no scaling, no mass), skip the right-hand side
  int warm;              // 1: start the solve from the current velocities, 0: from zero
  int has_ext;
};

template <typename R> struct IeLaunch {
  // max_width: the widest slice (slots per row)
  static void assemble(const DevMesh<R> &m, const TileDev &td, const DevParams &p, const IeDev &d, const IeAsm &a,
                       int max_width, hipStream_t st);
  static void start(const IeDev &d, hipStream_t st);
  // per-tet 12 x 12 stiffness of the solver tets ids[0 .. num) at positions x (planes), FP32 out
  static void tet_stiffness(const DevMesh<R> &m, const DevParams &p, const float *inst_mu, const float *inst_lam,
                            const R *x, int num, const int *ids, float *out, hipStream_t st);
  // one CG iteration = spmv, alpha, update, beta (+ the search-direction update), each a kernel;
  // e0, e1 (optional): events recording the start and end of the kernel's own dispatch
  static void spmv(const IeDev &d, hipStream_t st, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
  static void alpha(const IeDev &d, hipStream_t st);
  static void update(const IeDev &d, hipStream_t st, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
  static void beta(const IeDev &d, hipStream_t st);
  static void finish(const DevMesh<R> &m, const DevParams &p, float h, const IeDev &d, hipStream_t st);
  // projective dynamics: global matrix, momentum term + iterate start, local step, velocity update
  static void pd_matrix(const DevMesh<R> &m, const DevParams &p, float stiffness, const IeDev &d, hipStream_t st);
  static void pd_begin(const DevMesh<R> &m, const DevParams &p, const IeDev &d, double *srhs, int has_ext,
                       hipStream_t st);
  // inc_pos[4 t + corner] = position of (t, corner) in its node's incidence run
  static void pd_local(const DevMesh<R> &m, const DevParams &p, float stiffness, const int *inc_pos, const double *pos,
                       const double *srhs, double *out, hipStream_t st);
  static void pd_finish(const DevMesh<R> &m, const DevParams &p, float damping, const IeDev &d, hipStream_t st);
};

// Captured CG iterations for one matrix.
struct CgGraphs {
  hipGraphExec_t graph = nullptr;
  double key[2] = {0, 0};  // tol2, maxit the graph was built with
};

struct ImplicitState {
  int nslice = 0;
  long slots = 0;
  int max_width = 0;                                // slots per row of the widest slice
  IeDev dev{};                                      // backward Euler: 3x3 blocks
  IeDev pd{};                                       // projective dynamics: same layout and vectors, scalar values
  double *pd_s = nullptr;                           // projective dynamics: momentum term (3N)
  // The PD matrix depends only on the rest shape, the masses, the stiffness and the time step: it is
  // built once per (stiffness, h) with its own Jacobi diagonal and reused by later steps.
  bool pd_built = false;
  float pd_stiffness = 0.0f;
  double pd_h = 0.0;
  int *pd_inc_pos = nullptr;                        // projective dynamics: incidence slot of each (tet, corner)
  std::vector<int> slice_off_h, row_len_h, bcol_h;  // host copies (stiffness export)
  CgGraphs cg_ie, cg_pd;
  std::vector<hipEvent_t> ev;                       // timing events of the timed replays
  std::vector<void *> allocs;
};

void implicit_release(ImplicitState *S);
void implicit_params_changed(ImplicitState *S);

}  // namespace pbs
