/*
 * phybs.h -- C ABI of the MI355X (gfx950) elastic-body descent solver.
 *
 * This library replaces the GPU descent path of PhyBS (class CUDAInterface, FVM/CUDAInterface.h:20-102,
 * FVM/CUDAInterface.cu:482-1007) and exposes the CPU twin's material calls
 * (IsotropicNeohookeanMaterial / IsotropicMaterial) on the GPU.  All entry points take plain
 * pointers and sizes; host buffers are caller-owned and copied in/out; device memory is owned
 * by the solver handle.  Positions, velocities and forces cross the boundary as AoS float xyz
 * triples (the layout of the reference's Eigen MatrixXf(3,N), TetMesh.h:90); tets as 4 int32
 * per tet (MatrixXi(4,T), TetMesh.h:93).
 *
 * Errors: every function returns PBS_OK (0) or a negative PBS_ERR_* code; pbs_last_error()
 * returns the message of the calling thread's last failure.  A handle is not re-entrant: use
 * one handle per host thread.  Work runs on the handle's own HIP stream.
 */
#ifndef PHYBS_H
#define PHYBS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PBS_ABI_VERSION 1

#define PBS_OK 0
#define PBS_ERR_ARG -1   /* invalid argument (null pointer, bad size, bad enum) */
#define PBS_ERR_HIP -2   /* HIP runtime error (no device, allocation failure, launch failure) */
#define PBS_ERR_MESH -3  /* invalid mesh (index out of range, inverted or degenerate rest tet) */
#define PBS_ERR_IO -4    /* file could not be read or written */
#define PBS_ERR_STATE -5 /* call not valid in the handle's current state */

/* Semantics of the descent loop.  The reference ships two versions that differ. */
#define PBS_SEM_REF_GPU 0 /* CUDAInterface::doDescentOpt (CUDAInterface.cu:911-1007) -- default */
#define PBS_SEM_REF_CPU 1 /* DescentOptimize::doDescentOpt (DescentOptimize.cpp:91-188): gravity and
                             external forces counted twice in the gradient, energy clamp E>1e-6 */

#define PBS_FP32 32 /* FP32 state and per-tet arithmetic (the reference's precision) */
#define PBS_FP64 64 /* FP64 state and per-tet arithmetic, FP32 mesh data (needed for rel-grad <= 1e-4) */

#define PBS_MAT_NEOHOOKEAN 0 /* IsotropicNeohookeanMaterial (IsotropicNeohookeanMaterial.cpp) */
#define PBS_MAT_STRETCH 1    /* psi = sum_i mu (s_i - 1)^2 + lambda/2 (J - 1)^2 (no reference class) */

typedef struct pbs_solver pbs_solver;

/* Every constant the reference hard-codes, with its value as the default. */
typedef struct pbs_params {
  float dt;             /* time step, 0.03 (CUDAInterface.cu:644; DescentOptimize.cpp:5) */
  float grav[3];        /* gravity acceleration, (0,-9.8,0) (TetMesh.cpp:205) */
  float youngs;         /* E, 5e5 in TetMesh.cpp:185 -- default here 1e5 */
  float nu;             /* Poisson ratio 0.45 */
  float density;        /* 1000 */
  float fixed_mass;     /* mass given to constrained nodes, 1e8 (TetMesh.cpp:215) */
  float alpha0;         /* initial step length 0.1 (CUDAInterface.cu:914) */
  float alpha_decay;    /* 0.7 on each rollback (:948) */
  float alpha_min;      /* stop when alpha < 0.01 (:919) */
  int hd_refresh;       /* Hessian-diagonal refresh period 32 (:921) */
  int energy_check;     /* energy check period 8 (:942) */
  float damping;        /* velocity damping 0.99 (:637) */
  float init_momentum;  /* initial-guess momentum 0.2 (:491) */
  int profile_n;        /* Chebyshev profile length 3 (:732-738) */
  int profile_k[8];     /* 1, 7, 10 */
  float profile_rho[8]; /* 0.96, 0.991, 0.999 */
  int semantics;        /* PBS_SEM_* */
  int precision;        /* PBS_FP32 / PBS_FP64 */
  int material;         /* PBS_MAT_* */
  int iterations;       /* descent iterations per step, 96 (fvm.cpp:187) */
  double relgrad_tol;   /* 0: fixed iterations; > 0: iterate until |g_k| <= tol |g_0| (free DOFs, FP64) */
  int max_iterations;   /* iteration cap in rel-grad mode */
  int block_size;       /* threads per workgroup of the per-tet / per-node utility kernels, 0 = 256 */
  int tile_tets;        /* max tets per workgroup of the tile pass (16..4096); 0 = chosen from mesh size and
                           LDS; -1 = autotuned at creation (64, 128, 256, 512 timed on 32 passes) */
  int graph_passes;     /* descent passes per captured HIP graph replay: 0 = 32, -1 = no graph
                           (kernels launched one by one) */
  double mixed_switch;  /* PBS_FP64 with graphs only (no reference counterpart): while |g|/|g_0| of
                           the last graph replay is above this, the tile pass forms F in FP64 and
                           evaluates the per-tet forces in FP32 (neo-Hookean: from H = F - I) with
                           FP32 contributions; energies stay FP64; then FP64 to the end of the step
                           (also as soon as |g| <= 2e-4 |f|, the FP32 forces' rounding scale, or
                           after a rollback).  Convergence is only declared on FP64 passes.
                           0 = never (default). */
} pbs_params;

typedef struct pbs_step_stats {
  int passes;       /* loop bodies run (accepted iterations + rollbacks) */
  int iterations;   /* accepted iterations (the reference's final k) */
  int rollbacks;    /* line-search rollbacks (CUDAInterface.cu:946-958) */
  int converged;    /* rel-grad mode: tolerance met */
  int used_learned; /* learned warm start accepted (not rejected by the gradient test) */
  float alpha;      /* final step length */
  double energy;    /* last accepted incremental-potential value (FP64 sum) */
  double g0_norm;   /* |g| at the plain initial guess, free DOFs */
  double g_norm;    /* |g| at the last evaluated iterate */
  double gpu_ms;    /* device time of the step (HIP events on the handle's stream) */
  double g_learned; /* |g| at the learned initial guess (0 when the warm start is off) */
} pbs_step_stats;

/* ---- parameters / errors ---- */
int pbs_default_params(pbs_params *p);
const char *pbs_last_error(void);
int pbs_abi_version(void);

/* ---- solver lifetime ----
 * Replaces CUDAInterface::CUDAInterface(int num_nodes, const float *nodes, const float *restPoses,
 * const int *constraintsMask, int num_tets, const int *tets, float youngs, float nu, float density,
 * int csr_nnz, ...) (CUDAInterface.h:23-25, CUDAInterface.cu:641-740).  The CSR arguments of the
 * reference serve its backward-Euler path only and are not taken.  As in the reference, the
 * initial positions are the rest positions (reset() copies restPoses, CUDAInterface.cu:776-782);
 * `nodes` is accepted for signature parity and applied after the reset when non-NULL.
 * Rest tets must have det(Dm) > 0 (the AN convention of TetMesh.cpp:279-294). */
int pbs_create(int num_nodes, const float *nodes, const float *rest, const int *constraint_mask, int num_tets,
               const int *tets, const pbs_params *p, pbs_solver **out);
/* Batched solver (no reference counterpart): num_instances independent meshes stepped together
 * (each tile and node pass covers all of them in one launch, with a loop control -- line search,
 * Chebyshev, convergence -- per instance).  Instance i has inst_nodes[i] nodes and inst_tets[i]
 * tets, concatenated in nodes / rest / constraint_mask / tets in instance order; tets hold
 * indices into the concatenated nodes (inside the instance's range); youngs[i], nu[i] set the
 * instance's material (p->youngs / p->nu are not used).  Results equal those of separate solvers
 * with the same tile size.  The learned warm start needs a single-mesh solver. */
int pbs_create_batch(int num_instances, const int *inst_nodes, const int *inst_tets, const float *nodes,
                     const float *rest, const int *constraint_mask, const int *tets, const float *youngs,
                     const float *nu, const pbs_params *p, pbs_solver **out);
/* pbs_create_batch with a tile size per instance (inst_tile_tets[i] tets per tile for instance
 * i; p->tile_tets is not used).  Results equal those of separate solvers, each with its
 * instance's tile size.  No reference counterpart. */
int pbs_create_batch_tiled(int num_instances, const int *inst_nodes, const int *inst_tets, const float *nodes,
                           const float *rest, const int *constraint_mask, const int *tets, const float *youngs,
                           const float *nu, const int *inst_tile_tets, const pbs_params *p, pbs_solver **out);
int pbs_num_instances(const pbs_solver *s);
/* Loop statistics of one instance in the last pbs_step (gpu_ms and g_learned are 0).  pbs_step
 * itself reports, for a batch, passes = the most any instance ran, iterations / rollbacks /
 * converged summed, alpha the smallest, gradient norms combined as sqrt(sum of squares). */
int pbs_instance_stats(const pbs_solver *s, int instance, pbs_step_stats *st);
/* Tile sizes timed by the autotuner (pbs_params.tile_tets = -1) and their time per descent pass
 * in microseconds; *count = 0 when the solver was not autotuned.  For a batch of at least 8
 * instances the autotuner also tunes each class of instances (4 classes of equal count by tet
 * count) alone; an entry with tile_tets = -2 is then the batch with per-instance sizes (present
 * when the classes chose different sizes, or one size other than the fastest uniform one).  No
 * counterpart in the reference (it has no tile pass). */
int pbs_autotune_report(const pbs_solver *s, int max, int *tile_tets, double *us_per_pass, int *count);
/* Tile size of each instance (pbs_num_instances values); after autotuning, a batch's instances
 * may have different sizes.  No counterpart in the reference. */
int pbs_instance_tile_tets(const pbs_solver *s, int *tile_tets);
/* ~CUDAInterface (CUDAInterface.cu:742-769). */
void pbs_destroy(pbs_solver *s);
/* Change solver parameters (material constants and gravity re-run the setup kernels). */
int pbs_set_params(pbs_solver *s, const pbs_params *p);
int pbs_get_params(const pbs_solver *s, pbs_params *p);
int pbs_sizes(const pbs_solver *s, int *num_nodes, int *num_tets);

/* Rest-shape data computed by the GPU setup kernels (compute_DmInv_ANs_Mass and the ctor,
 * CUDAInterface.cu:273-325, :704-725): dminv 9T (row-major Dm^-1 per tet), vol T, mass N
 * (constrained nodes hold fixed_mass), gforce 3N AoS (mass before the override * grav).
 * Any pointer may be NULL. */
int pbs_get_rest_data(pbs_solver *s, float *dminv, float *vol, float *mass, float *gforce);

/* ---- state ---- */
/* CUDAInterface::reset (CUDAInterface.cu:776-782): positions := rest, velocities := 0. */
int pbs_reset(pbs_solver *s);
/* CUDAInterface::updateNodePositions (CUDAInterface.cu:771-774). */
int pbs_set_positions(pbs_solver *s, const float *xyz);
int pbs_get_positions(pbs_solver *s, float *xyz);
int pbs_get_positions_f64(pbs_solver *s, double *xyz);
int pbs_set_velocities(pbs_solver *s, const float *v, const float *v_last);
int pbs_get_velocities(pbs_solver *s, float *v, float *v_last);
/* Binary snapshot of the state a step starts from -- positions, velocities and previous
 * velocities (x, v, v_last) at the solver's precision -- and its restore into a solver built on
 * the same meshes (same node counts per instance and precision; otherwise PBS_ERR_ARG and the
 * state is left unchanged).  The reference checkpoints positions only, as a .node file
 * (TetMesh.cpp:558-570, restored through CUDAInterface::updateNodePositions, .cu:771-774); a
 * restored FP64 solver continues bitwise as the saved one would have. */
int pbs_state_save(pbs_solver *s, const char *path);
int pbs_state_load(pbs_solver *s, const char *path);

/* ---- the hot path ----
 * One implicit-Euler step by Jacobi-preconditioned descent with Chebyshev acceleration and the
 * energy rollback line search.  ext: 3N AoS external forces or NULL; out: 3N AoS positions
 * after the step or NULL; st: optional statistics. */
int pbs_step(pbs_solver *s, const float *ext, float *out, pbs_step_stats *st);
/* CUDAInterface::doDescentOpt(float h, int iterations, float *hostExternForces, float *hostNode)
 * (CUDAInterface.h:35, CUDAInterface.cu:911-1007).  As in the reference, h is ignored (the
 * handle's dt is used) and ext must not be NULL. */
int pbs_do_descent_opt(pbs_solver *s, float h, int iterations, const float *ext, float *out);

/* ---- material calls on the GPU ----
 * IsotropicNeohookeanMaterial::computeInnerForceFromPos / computeElasticEnergyFromPos
 * (IsotropicNeohookeanMaterial.cpp:93-131) at positions pos (3N AoS).  Outputs (any may be NULL):
 * tet_forces 12T (forces on the tet's 4 nodes, xyz each), node_forces 3N (elastic only: the
 * reference adds gravity, external and fixed-node spring forces there, TetMesh.cpp:580-586),
 * tet_energy T (psi * vol with the semantics' clamp). */
int pbs_compute_forces(pbs_solver *s, const float *pos, float *tet_forces, float *node_forces, float *tet_energy);
/* Diagonal of IsotropicMaterial::computeGlobalStiffnessMatrixFromPos (IsotropicMaterial.cpp:646-682),
 * i.e. of K = -df/dx: node_hd 3N, tet_hd 12T (either may be NULL). */
int pbs_compute_hessian_diag(pbs_solver *s, const float *pos, float *node_hd, float *tet_hd);
/* The per-node elastic forces and stiffness diagonal as the descent pass itself forms them: one
 * launch of the tile kernel at pos (3N AoS) -- the per-tet half of computeForceKd
 * (CUDAInterface.cu:495-568) -- then each node's owned partial plus its halo partials, in the
 * order the node pass adds them.  node_forces 3N (elastic only, as pbs_compute_forces), node_hd
 * 3N (NULL: no Hessian phase), energy: the pass's FP64 elastic energy (may be NULL).
 * Single-mesh handles only. */
int pbs_tile_forces(pbs_solver *s, const float *pos, double *node_forces, double *node_hd, double *energy);

/* ---- measurement ---- */
/* Average device time per launch (ms) of each kernel of the last pbs_step calls, measured with
 * HIP events on the handle's stream: names[i] / avg_ms[i] / launches[i] for i < *count.  The
 * descent-pass kernels and the CG solves' SpMV and vector update are timed by events attached to
 * their own dispatch (the interval a kernel trace reports); step start, warm start, end, the
 * backward-Euler assembly and the PD local step by events recorded around them. */
int pbs_kernel_times(pbs_solver *s, int max, const char **names, double *avg_ms, long *launches, int *count);
int pbs_reset_kernel_times(pbs_solver *s);
/* period 0: no timing (default); 1: every descent-pass kernel is timed; n > 1: the kernels of
 * every n-th graph replay (every n-th pass without graphs) are timed, the timed replay's passes
 * being launched one by one instead of from the graph.  Launches that run after the loop has
 * stopped within a replay are not counted.  Any period > 0 times every CG replay. */
int pbs_set_profiling(pbs_solver *s, int period);

/* ---- backward Euler (the reference's GPU implicit path, fvm.cpp:179-183) ---- */
typedef struct pbs_implicit_params {
  float damping_alpha;   /* 0.02: mass damping, M (1 + alpha h) (m_dumpingAlpha, CUDAInterface.cu:644) */
  float damping_beta;    /* 0.02: stiffness damping, K (h + beta) h (m_dumpingBelta) */
  double cg_tolerance;   /* 1e-6: stop at |D^-1/2 (b - A v)| <= tol |D^-1/2 b|, D = diag(A) (FP64) */
  int cg_max_iterations; /* 10000 (the cap of CUDALinearSolvers::conjugateGradient) */
  int warm_start;        /* 1: CG starts from the current velocities (as that solver does), 0: zero */
} pbs_implicit_params;
typedef struct pbs_implicit_stats {
  int cg_iterations;
  int converged;         /* 0: stopped at cg_max_iterations (the approximate solution is applied) */
  double residual0, residual, rhs_norm; /* the Jacobi-scaled norms of b - A v0, b - A v, b */
  float gpu_ms;          /* device time of the step */
} pbs_implicit_stats;
int pbs_implicit_default_params(pbs_implicit_params *p);
/* One backward-Euler step (CUDAInterface::computeInnerforces + computeGlobalStiffnessMatrix +
 * doBackEuler, CUDAInterface.cu:784-884): with f the elastic forces plus gravity (fixed nodes: the
 * spring 1e8 (rest - x) instead) and ext (3N AoS, may be NULL: the reference has none),
 *   [(1 + alpha h) M + (h + beta) h K(x)] v' = h f + M v,   x' = x + h v'   (all nodes),
 * solved by Jacobi-preconditioned CG on the device (the reference factorises with cuSOLVER's
 * sparse Cholesky).  out (3N AoS, may be NULL) receives x'.  A system that is not positive
 * definite leaves the state unchanged and returns PBS_ERR_STATE.  Not available on batches. */
int pbs_implicit_step(pbs_solver *s, const pbs_implicit_params *p, const float *ext, float *out,
                      pbs_implicit_stats *st);
/* K(x) = -df/dx at the current positions in the reference's CSR layout (computeGlobalK,
 * CUDAInterface.cu:234-271; pattern = pbs_csr_pattern's rowptr/col, nnz values). */
int pbs_stiffness_matrix(pbs_solver *s, int64_t nnz, double *csr_val);
/* IsotropicMaterial::computeStiffnessMatrixFromPos(tetID, pos) (IsotropicMaterial.cpp:503-552) for
 * num tets at once: K (144 FP32 values per listed tet, row-major 12 x 12, row 3a + d and column
 * 3b + e for the tet's corners a, b in its own corner order and axes d, e) = -df/dx of the tet's
 * elastic forces at positions pos (3N AoS).  Single-mesh handles only. */
int pbs_tet_stiffness(pbs_solver *s, const float *pos, int num, const int *tet_ids, float *K);
/* No counterpart in the reference.  pbs_tet_stiffness without any device allocation, for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops): its device buffers
 * (tet ids, matrices) are a fixed 4.75 MB array per GPU, held until the process ends.  At most
 * 8192 tets per call (PBS_ERR_ARG beyond).  Calls that overlap on one GPU share that array and
 * race; results equal pbs_tet_stiffness's. */
int pbs_tet_stiffness_noalloc(pbs_solver *s, const float *pos, int num, const int *tet_ids, float *K);

/* ---- projective dynamics (ProjDynamic::doProjDynamics, ProjDynamic.cpp:442-479) ---- */
typedef struct pbs_pd_params {
  float stiffness;       /* 1e6: constraint weight per unit volume (m_stiffness, ProjDynamic.cpp:4) */
  int iterations;        /* 5 local/global rounds (m_iterations) */
  float damping;         /* 0.99: v = damping (x' - x) / dt (:477) */
  double cg_tolerance;   /* 1e-10: each global solve, in the Jacobi-scaled norm, whose |D^-1/2 b|
                            the fixed nodes' 1e8-mass rows dominate (the reference factorises
                            once with Pardiso LDLT) */
  int cg_max_iterations; /* 10000 per global solve */
} pbs_pd_params;
typedef struct pbs_pd_stats {
  int cg_iterations;     /* summed over the rounds */
  int converged;         /* every global solve met the tolerance */
  float gpu_ms;
} pbs_pd_stats;
int pbs_pd_default_params(pbs_pd_params *p);
/* One projective-dynamics step with the handle's dt, mass and gravity: s = M (x + dt v)/dt^2 +
 * f_ext (free nodes; f_ext = gravity + ext, ext may be NULL), then `iterations` rounds of the
 * local step (per-tet rotation R of F, see pbs_pd_project) and the global solve
 * [M/dt^2 + sum_k w_k Ac_k^T Ac_k] x = s + projections; v' = damping (x' - x)/dt, x = x'. */
int pbs_pd_step(pbs_solver *s, const pbs_pd_params *p, const float *ext, float *out, pbs_pd_stats *st);
/* Local step alone (projectLocalConstraints2, ProjDynamic.cpp:223-308): for positions (3N AoS),
 * proj (3N AoS) = sum over each node's tets k of w_k R_k c_i, w_k = stiffness * vol_k, c_i the
 * node's shape vector and R_k = U V^T of F_k with the reference's sign fixes (:359-368). */
int pbs_pd_project(pbs_solver *s, float stiffness, const float *positions, double *proj);

/* ---- host mesh utilities (no GPU work) ---- */
/* Bar of nx*ny*nz cubes of size edge, 6 tets per cube (Freudenthal split, det(Dm) > 0).
 * nodes: 3*(nx+1)(ny+1)(nz+1) floats, tets: 24*nx*ny*nz ints; node id = i + (nx+1)(j + (ny+1)k). */
int pbs_mesh_bar(int nx, int ny, int nz, float edge, float *nodes, int *tets);
/* Twist about the bar's z axis (theta = twist_deg * z / (nz*edge)) plus uniform +-jitter*edge
 * SplitMix64(seed) noise per coordinate, computed in FP64 and rounded (in place, 3N AoS). */
int pbs_mesh_twist(int nx, int ny, int nz, float edge, double twist_deg, double jitter, uint64_t seed,
                   int num_nodes, float *nodes);
/* TetGen .node/.ele/.face (TetMesh.cpp:14-151).  Counts first, then load; positions are
 * multiplied by scale (the reference uses 0.1, TetMesh.cpp:38).  faces may be NULL. */
int pbs_tetgen_count(const char *prefix, int *num_nodes, int *num_tets, int *num_faces);
int pbs_tetgen_load(const char *prefix, float scale, float *nodes, int *tets, int *faces);
/* One file at a time, as TetMesh::initNodesFromFile / initTetsFromFile / initFacesFromFile
 * (TetMesh.h:20-22; TetMesh.cpp:14-56, 96-125, 127-151) read them: the count is returned in
 * *num_*; the array (3 floats per node, times scale; 4 or 3 ints per tet or face) is filled when
 * non-NULL.  The reference's hard-coded constraint rule is not applied here (phybs.hpp's TetMeshT
 * applies it). */
int pbs_tetgen_read_nodes(const char *path, float scale, int *num_nodes, float *nodes);
int pbs_tetgen_read_tets(const char *path, int *num_tets, int *tets);
int pbs_tetgen_read_faces(const char *path, int *num_faces, int *faces);
/* .node export / import (TetMesh.cpp:558-570 / :70-94, no scaling on import). */
int pbs_node_write(const char *path, int num_nodes, const float *nodes);
int pbs_node_read(const char *path, int num_nodes, float *nodes);
/* Stiffness pattern of IsotropicMaterial::allocateGlobalStiffnessMatrix (IsotropicMaterial.cpp:554-644):
 * returns nnz (>= 0) or an error; when rowptr/colidx are non-NULL fills rowptr (3N+1), colidx
 * (nnz), diag_idx (3N, getDiagonalIdx) and kid (48T, getKIDinCSRval). */
int64_t pbs_csr_pattern(int num_nodes, int num_tets, const int *tets, int *rowptr, int *colidx, int *diag_idx,
                        int *kid);
/* Tet incidence per node in ascending tet order (ProjDynamic::initLocalProjection,
 * ProjDynamic.cpp:25-45): offsets N+1, entries 4T = 12*tet + 3*local. */
int pbs_incidence(int num_nodes, int num_tets, const int *tets, int *offsets, int *entries);

/* ---- interactive external forces (host; the runtime input of doDescentOpt) ---- */
/* Ring of each boundary face (TetMesh::computeBoundfaceRingIndices, TetMesh.cpp:229-276): its 3
 * nodes in face order, then the other nodes sharing a boundary edge with any of them, ascending.
 * Returns the total length; fills offsets (F+1) and ring when non-NULL. */
int64_t pbs_face_ring(int num_nodes, int num_faces, const int *faces, int64_t *offsets, int *ring);
/* Face hit by a ray (TetMesh::pickFacebyRay, TetMesh.cpp:393-418): *face = the face facing the ray
 * whose centre is nearest to it, within the length of face 0's first edge; -1 if none. */
int pbs_pick_face(int num_nodes, const float *nodes, int num_faces, const int *faces, const float *origin,
                  const float *direction, int *face);
/* External forces of dragging a face (TetMesh::dragFaceRing, TetMesh.cpp:446-470): factor =
 * min(|drag| / |p1 - p0 of face 0| * 100, 800); factor * mass * drag direction on the face's nodes,
 * half of it on the rest of its ring; ext (3N AoS) is overwritten. */
int pbs_drag_face_ring(int num_nodes, const float *nodes, const float *mass, int num_faces, const int *faces,
                       const int64_t *ring_offsets, const int *ring, int face, const float *drag, float *ext);

/* ---- learned warm start (no counterpart in the reference, which has no learned component) ----
 * A per-vertex network (18 features -> 64 -> 64 -> 64 -> 3, SiLU; BF16 weights on the matrix
 * cores, FP32 accumulation) predicts the displacement of the converged step from the inertial
 * target y = x + h v + h^2 g.  Features per free node at the step start: u = x - X, h v,
 * s = h v + h^2 g and the 1-ring means of (u, h v, s) minus the node's own.  With the warm start
 * enabled, pbs_step evaluates the gradient at the plain and at the learned guess and starts from
 * the learned one only if its gradient norm is not larger; the plain norm remains g0_norm (the
 * rel-grad denominator).  Parameters: PBS_MLP_PARAMS floats, packed W1 (18x64, in x out), b1,
 * W2 (64x64), b2, W3, b3, W4 (64x3), b4; PBS_MLP_SCALES scales (6 feature groups, output). */
#define PBS_MLP_PARAMS 9731
#define PBS_MLP_SCALES 7
int pbs_learn_set_model(pbs_solver *s, const float *params, const float *scales);
int pbs_learn_get_model(pbs_solver *s, float *params, float *scales);
/* Glorot-uniform weights from seed, zero biases (scales unchanged). */
int pbs_learn_init_model(pbs_solver *s, uint64_t seed);
int pbs_learn_enable(pbs_solver *s, int on);
/* How the learned guess is used: the correction d = learned - plain is smoothed by `sweeps`
 * averaging sweeps over the node graph (d <- d/2 + (1-ring mean of d)/2, fixed nodes at 0), then
 * the candidates plain + beta_j d (j < nbeta, 1..16 steps >= 0) are each evaluated with one
 * gradient pass and the best is used if its gradient norm is not above the plain guess's.
 * Defaults: 0 sweeps, one step of 1. */
int pbs_learn_set_options(pbs_solver *s, int sweeps, int nbeta, const float *betas);
/* Starts collecting (features, converged x - y) of every free node from each pbs_step, up to
 * capacity samples (0 stops and frees); collection stops when the next step would not fit.
 * Samples are stored step after step, the free nodes of a step in the caller's node order. */
int pbs_learn_collect(pbs_solver *s, int64_t capacity);
/* The same, keeping from the k-th collected step only the free nodes whose rank r among the free
 * nodes (caller's order) has r % stride == k % stride: every node once per stride steps. */
int pbs_learn_collect_every(pbs_solver *s, int64_t capacity, int stride);
int64_t pbs_learn_samples(pbs_solver *s);
int pbs_learn_get_samples(pbs_solver *s, int64_t start, int64_t n, float *features, float *targets);
/* Appends n samples (18 features, 3 targets each) to the collected ones, e.g. those another
 * handle collected on another mesh, so that one network trains on several configurations;
 * fails if they do not fit the capacity given to pbs_learn_collect*. */
int pbs_learn_add_samples(pbs_solver *s, int64_t n, const float *features, const float *targets);
/* Scales = RMS of each feature group and of the targets over the collected samples. */
int pbs_learn_fit_scales(pbs_solver *s);
/* Adam (lr, beta 0.9/0.999) on the mean squared error of the normalised outputs, mini-batches of
 * `batch` shuffled samples, on the GPU; loss_per_epoch (may be NULL) gets each epoch's mean loss. */
int pbs_learn_train(pbs_solver *s, int epochs, int batch, float lr, uint64_t seed, float *loss_per_epoch);
/* Network outputs (normalised, 3 per sample) for n feature rows (18 each). */
int pbs_learn_forward(pbs_solver *s, int n, const float *features, float *out);
/* Mean loss and gradient (PBS_MLP_PARAMS) over n samples, without updating the model. */
int pbs_learn_loss_grad(pbs_solver *s, int n, const float *features, const float *targets, double *loss, float *grad);
/* Features of the current state (N x 18, caller's node order; zero rows for fixed nodes). */
int pbs_learn_features(pbs_solver *s, float *features);
/* The learned and the plain initial guess of the next step (3N AoS doubles each, either may be
 * NULL); the solver state is not changed. */
int pbs_learn_guess(pbs_solver *s, double *learned, double *plain);

#ifdef __cplusplus
}
#endif
#endif
