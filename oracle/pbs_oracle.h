/*
 * pbs_oracle.h -- CPU restatement of PhyBS's elastic-body descent path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP solver in
 * phybs_amd/: only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it.  No product code links or calls it.
 *
 * Every function restates the reference's algorithm (file:line cited at each
 * definition in pbs_oracle.c / oracle_impl.h).  Two precisions are built from one
 * source: *_f32 keeps the reference's FP32 storage and arithmetic, *_f64 evaluates
 * the same formulas in FP64.  Sums that decide the line search are FP64 in both.
 *
 * Parity pins (what checks each function) are listed in the header of
 * pbs_oracle.c.  Functions without an independent pin say "parity unpinned".
 */
#ifndef PBS_ORACLE_H
#define PBS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Semantics of the two reference descent paths (they differ; see oracle_impl.h). */
#define ORC_REF_GPU 0 /* CUDAInterface::doDescentOpt, CUDAInterface.cu:911-1007 */
#define ORC_REF_CPU 1 /* DescentOptimize::doDescentOpt, DescentOptimize.cpp:91-188 */

/* Energy clamp variants: none, GPU max(E,0) (CUDAInterface.cu:549),
 * CPU E>1e-6?E:0 (IsotropicNeohookeanMaterial.cpp:129). */
#define ORC_CLAMP_NONE 0
#define ORC_CLAMP_GPU 1
#define ORC_CLAMP_CPU 2

#define ORC_MAT_NEOHOOKEAN 0 /* IsotropicNeohookeanMaterial */
#define ORC_MAT_STRETCH 1    /* psi = sum mu (s_i-1)^2 + lambda/2 (J-1)^2 (no reference counterpart) */

typedef struct orc_params {
  float dt;             /* m_timestep 0.03 (CUDAInterface.cu:644) */
  float grav[3];        /* (0,-9.8,0) (TetMesh.cpp:205) */
  float alpha0;         /* 0.1 (CUDAInterface.cu:914) */
  float alpha_decay;    /* 0.7 (:948) */
  float alpha_min;      /* 0.01 (:919) */
  int hd_refresh;       /* 32 (:921) */
  int energy_check;     /* 8 (:942) */
  float damping;        /* 0.99 (:637) */
  float init_momentum;  /* 0.2 (:491) */
  int profile_n;        /* 3 (:732-738) */
  int profile_k[8];
  float profile_rho[8];
  int semantics;        /* ORC_REF_GPU / ORC_REF_CPU */
  int iterations;       /* per step (fvm.cpp:187 passes 96) */
  double relgrad_tol;   /* 0: fixed-iteration mode; >0: stop when |g_k|/|g_0| <= tol */
  int trace_cap;        /* number of per-pass trace records kept (0 = none) */
  int gather_order;     /* per-node force / Hessian-diagonal sum: 0 = ascending tet id (the CPU
                           path's tet loop order), 1 = descending.  The GPU path sums with float
                           atomics (CUDAInterface.cu:526-541), i.e. in no fixed order; order 1
                           measures how far FP32 results move with the summation order alone. */
} orc_params;

typedef struct orc_stats {
  int passes;           /* loop bodies executed (accepted iterations + rollbacks) */
  int iterations;       /* final value of k (accepted iterations) */
  int rollbacks;
  int converged;        /* rel-grad mode: 1 when the tolerance was met */
  float alpha;          /* final step length */
  double energy;        /* last checked energy */
  double g0_norm;       /* |g(x_0)| over free DOFs */
  double g_norm;        /* |g| of the last evaluated iterate */
  int used_learned;     /* the step started from the caller's guess (its gradient was not larger) */
} orc_stats;

/* One record per loop pass (for tests of the control flow). */
typedef struct orc_trace {
  int k, rollback, hd_refresh, checked;
  float alpha, omega;
  double energy, g_norm;
} orc_trace;

/* ---- material constants (IsotropicNeohookeanMaterial.cpp:15-16) ---- */
void orc_lame(float E, float nu, float *mu, float *lambda);

/* ---- rest-shape precompute (TetMesh.cpp:183-216, 279-294) ---- */
/* rest: 3N AoS; tets: 4T; mask: N (1 = fixed).  Outputs: dminv 9T row-major,
 * vol T, an 9T row-major (columns = area vectors of local nodes 1..3),
 * mass N (fixed nodes := fixed_mass), gforce 3N (= mass before the fixed override * grav). */
void orc_precompute(int nn, const float *rest, int nt, const int *tets, const int *mask,
                    float density, const float grav[3], float fixed_mass,
                    float *dminv, float *vol, float *an, float *mass, float *gforce);

/* ---- per-tet force / energy (IsotropicNeohookeanMaterial.cpp:93-131; CUDAInterface.cu:495-551) ---- */
void orc_tet_forces_f32(int nt, const int *tets, const float *x, const float *dminv, const float *an,
                        const float *vol, float mu, float lam, int material, int clamp,
                        float *tetf, float *tete);
void orc_tet_forces_f64(int nt, const int *tets, const double *x, const float *dminv, const float *an,
                        const float *vol, float mu, float lam, int material, int clamp,
                        double *tetf, double *tete);

/* ---- 3x3 SVD with the reference's sign conventions (IsotropicMaterial.cpp:16-60) ---- */
void orc_svd3_f32(const float *F, float *U, float *S, float *V);
void orc_svd3_f64(const double *F, double *U, double *S, double *V);

/* ---- dP/dF by Xu et al. 2015 with the reference perturbation (IsotropicMaterial.cpp:856-1020) ----
 * out[81]: row-major over (P index, F index) as the GPU stores it (CUDAInterface.cu:173). */
void orc_dpdf_f32(const float *F, float mu, float lam, int material, float *out);
void orc_dpdf_f64(const double *F, float mu, float lam, int material, double *out);

/* ---- Hessian (stiffness) diagonal (IsotropicMaterial.cpp:503-552, 646-682) ----
 * hd: 3N, per-node diagonal of K = -df/dx summed over incident tets (ascending tet order).
 * tethd (optional, 12T): per-tet contributions. */
void orc_hessian_diag_f32(int nn, int nt, const int *tets, const float *x, const float *dminv,
                          const float *an, float mu, float lam, int material, float *hd, float *tethd);
void orc_hessian_diag_f64(int nn, int nt, const int *tets, const double *x, const float *dminv,
                          const float *an, float mu, float lam, int material, double *hd, double *tethd);

/* ---- backward-Euler linear system (fvm.cpp:179-183; CUDAInterface.cu:234-271, 327-452, 819-852) ----
 * kval: K = -df/dx in the CSR layout of orc_csr_pattern (nnz values, via kid); aval: the system
 * matrix (h + beta) h K + (1 + alpha h) M on the same pattern; b (3N): h f + M v with
 * f = elastic (SVD route) + gravity, fixed nodes 1e8 (rest - x).  rest, gforce: 3N AoS;
 * x, v, ext: 3N AoS (ext may be NULL). */
void orc_implicit_system_f32(int nn, const float *rest, int nt, const int *tets, const int *mask,
                             const float *dminv, const float *an, const float *mass, const float *gforce,
                             float mu, float lam, int material, const float *x, const float *v, const float *ext,
                             float h, float alpha, float beta, long nnz, const int *kid, const int *diag_idx,
                             double *kval, double *aval, double *b);
void orc_implicit_system_f64(int nn, const float *rest, int nt, const int *tets, const int *mask,
                             const float *dminv, const float *an, const float *mass, const float *gforce,
                             float mu, float lam, int material, const double *x, const double *v,
                             const double *ext, float h, float alpha, float beta, long nnz, const int *kid,
                             const int *diag_idx, double *kval, double *aval, double *b);

/* ---- projective dynamics (ProjDynamic.cpp:46-96, 223-308, 359-368) ----
 * rotation: R = U V^T with the reference's sign fixes; project: out = s + per-node sums of
 * w R c_i (s may be NULL), x/s/out 3N AoS; matrix: M/t^2 + sum w Ac^T Ac on orc_csr_pattern. */
void orc_pd_rotation_f32(const float *F, float *Rot);
void orc_pd_rotation_f64(const double *F, double *Rot);
void orc_pd_project_f32(int nn, int nt, const int *tets, const float *x, const float *dminv, const float *vol,
                        float stiffness, const float *s, float *out);
void orc_pd_project_f64(int nn, int nt, const int *tets, const double *x, const float *dminv, const float *vol,
                        float stiffness, const double *s, double *out);
void orc_pd_matrix_f32(int nn, int nt, const int *tets, const float *dminv, const float *vol, const float *mass,
                       float stiffness, float t, long nnz, const int *kid, const int *diag_idx, double *gval);
void orc_pd_matrix_f64(int nn, int nt, const int *tets, const float *dminv, const float *vol, const float *mass,
                       float stiffness, float t, long nnz, const int *kid, const int *diag_idx, double *gval);

/* ---- adjacency (IsotropicMaterial.cpp:554-644, IsotropicMaterial.h:120-138; ProjDynamic.cpp:25-45) ---- */
/* Pattern of the 3N x 3N stiffness matrix: returns nnz; rowptr (3N+1) and colidx (nnz) are
 * written when non-NULL (call once with NULLs to size).  diag_idx (3N) and kid (48T) as the
 * reference's getDiagonalIdx()/getKIDinCSRval(). */
long orc_csr_pattern(int nn, int nt, const int *tets, int *rowptr, int *colidx, int *diag_idx, int *kid);
/* Tet incidence per node: offsets (N+1), entries (4T) = 12*tet + 3*local, ascending by tet. */
void orc_incidence(int nn, int nt, const int *tets, int *offsets, int *entries);

/* ---- descent solver (CUDAInterface.cu:482-1007; DescentOptimize.cpp:29-188) ---- */
typedef struct orc_solver orc_solver;
void orc_default_params(orc_params *p);
/* precision: 32 or 64. */
orc_solver *orc_create(int nn, const float *rest, int nt, const int *tets, const int *mask,
                       float E, float nu, float density, float fixed_mass, int material, int precision,
                       const orc_params *p);
void orc_destroy(orc_solver *s);
void orc_set_params(orc_solver *s, const orc_params *p);
void orc_set_threads(orc_solver *s, int nthreads);
/* State I/O in FP64 (rounded to the solver precision on input). */
void orc_set_state(orc_solver *s, const double *x, const double *v, const double *v_last);
void orc_get_state(orc_solver *s, double *x, double *v, double *v_last);
/* ext: 3N AoS or NULL.  x0_override: 3N initial guess replacing the inertial guess
 * (NULL = the reference's guess).  trace: params.trace_cap records or NULL. */
int orc_step(orc_solver *s, const double *ext, const double *x0_override, orc_stats *st, orc_trace *trace);
/* Gradient of the incremental potential at x for the current step (after orc_step set x~). */
double orc_gradient_norm(orc_solver *s, const double *x);

#ifdef __cplusplus
}
#endif
#endif
