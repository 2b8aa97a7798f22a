/*
 * pbs_oracle.c -- CPU restatement of PhyBS's descent path (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library,
 * and only as the checker / CPU baseline.  The product (phybs_amd/) never links it.
 *
 * The reference's CPU path (the FVM .cpp files) is not buildable in this image without stand-in
 * headers (it includes <Eigen\Dense> with a Windows path separator, <Eigen/svd>, Qt and
 * QGLViewer/GL headers), so every function here is pinned by independent checks in tests/:
 *
 *   orc_lame            closed form E, nu -> mu, lambda including the lambda = 2mu/(1-2nu) quirk
 *                       (test_oracle.py::test_lame)
 *   orc_precompute      vol = edge^3/6 for grid tets, sum of masses = rho V, AN = -vol Dm^-T,
 *                       Dm^-1 Dm = I, gravity = m g (test_oracle.py::test_precompute_*)
 *   orc_tet_forces      P(I) = 0, P(sI) closed form, P(RF) = R P(F), sum f = 0, sum x x f = 0,
 *                       f = -dE/dx by FP64 central differences of the energy, stretch material
 *                       against an independent numpy P (test_oracle.py::test_forces_*)
 *   orc_svd3            U S V^T = F, det U = det V = +1, orthogonality, |S| = numpy singular
 *                       values, sign conventions (test_oracle.py::test_svd_*)
 *   orc_dpdf            central differences of P, closed-form neo-Hookean dP/dF written in
 *                       index form in numpy (test_oracle.py::test_dpdf_*)
 *   orc_hessian_diag    diagonal of -df/dx by central differences of the forces
 *                       (test_oracle.py::test_hessian_diag_fd)
 *   orc_implicit_system K = -df/dx by central differences of its own right-hand side (whose
 *                       forces come from the SVD route, pinned against the closed-form forces),
 *                       translation null space K t = 0, symmetry, the system matrix against
 *                       its definition (test_oracle.py::test_implicit_*)
 *   orc_pd_rotation     orthogonal with det +1; the polar factor (scipy.linalg.polar) when det F > 0;
 *                       numpy's SVD with the reference's sign fixes when det F < 0
 *                       (test_oracle.py::test_pd_rotation_*)
 *   orc_pd_project      the gradient of sum_k w_k <F_k(x), R_k> with R_k frozen, by central
 *                       differences (test_oracle.py::test_pd_project_fd)
 *   orc_pd_matrix       the Hessian of sum_k w_k |F_k(x)|^2 / 2 by differences of its gradient,
 *                       plus M / t^2 (test_oracle.py::test_pd_matrix_fd)
 *   orc_csr_pattern     brute force from the tet list: row pattern, diagonal index, kID
 *                       (test_mesh.py::test_csr_pattern_bit_exact)
 *   orc_incidence       brute force (test_mesh.py::test_incidence_bit_exact)
 *   orc_step            mass-only problem with the closed-form minimiser, REF_CPU double-counted
 *                       gravity minimiser, convergence to an independent scipy minimisation
 *                       of the incremental potential on small meshes, energy monotone on
 *                       accepted checks, rollback rules, Chebyshev omega limit
 *                       (test_oracle.py::test_step_*)
 *
 * Intentional departures from the reference's arithmetic (all below rounding of the
 * tolerances used): the unconstrained SVD is one-sided Jacobi in FP64 (the reference calls
 * Eigen::JacobiSVD<Matrix3f> / McAdams), line-search energies are summed in FP64 (the
 * reference sums FP32 with atomics + thrust::reduce), per-node forces are gathered in
 * ascending tet order (the GPU reference uses float atomics; orc_params.gather_order = 1 sums in
 * descending order, to size the FP32 order sensitivity, test_oracle.py::test_gather_order_*).
 */
#include "pbs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct orc_solver {
  int nn, nt, material, precision;
  int *tets, *mask;
  float *dminv, *an, *vol, *mass, *gforce;
  float mu, lam;
  orc_params p;
  int *inc_off, *inc;
  void *state;
  float rho, omega;
  double e_old;
};

/* Eigen's unitOrthogonal for a 3-vector: (-y, x, 0)/n unless x and y are ~0, else (0, -z, y)/n. */
static void unit_orthogonal(double x, double y, double z, double *o) {
  if (!(fabs(x) <= 1e-12 * fabs(z) && fabs(y) <= 1e-12 * fabs(z))) {
    double n = sqrt(x * x + y * y);
    o[0] = -y / n; o[1] = x / n; o[2] = 0;
  } else {
    double n = sqrt(y * y + z * z);
    o[0] = 0; o[1] = -z / n; o[2] = y / n;
  }
}

/* One-sided (Hestenes) Jacobi SVD of a 3x3 row-major matrix in FP64: columns of W = F V are
 * orthogonalised by plane rotations; S = column norms sorted descending (nonnegative);
 * V orthogonal (det may be -1; orc_svd3 applies the reference's sign conventions). */
static void svd3_core(const double *F, double *S, double *V) {
  double W[9];
  memcpy(W, F, sizeof(W));
  for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
  static const int pr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int sweep = 0; sweep < 64; ++sweep) {
    int rotated = 0;
    for (int q = 0; q < 3; ++q) {
      int a = pr[q][0], b = pr[q][1];
      double al = 0, be = 0, ga = 0;
      for (int r = 0; r < 3; ++r) {
        al += W[3 * r + a] * W[3 * r + a];
        be += W[3 * r + b] * W[3 * r + b];
        ga += W[3 * r + a] * W[3 * r + b];
      }
      if (ga == 0.0 || fabs(ga) <= 1e-17 * sqrt(al * be)) continue;
      rotated = 1;
      double z = (be - al) / (2.0 * ga);
      double t = (z >= 0 ? 1.0 : -1.0) / (fabs(z) + sqrt(1.0 + z * z));
      double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
      for (int r = 0; r < 3; ++r) {
        double wa = W[3 * r + a], wb = W[3 * r + b];
        W[3 * r + a] = c * wa - s * wb;
        W[3 * r + b] = s * wa + c * wb;
        double va = V[3 * r + a], vb = V[3 * r + b];
        V[3 * r + a] = c * va - s * vb;
        V[3 * r + b] = s * va + c * vb;
      }
    }
    if (!rotated) break;
  }
  double sv[3];
  for (int c = 0; c < 3; ++c) sv[c] = sqrt(W[c] * W[c] + W[3 + c] * W[3 + c] + W[6 + c] * W[6 + c]);
  int idx[3] = {0, 1, 2};
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j)
      if (sv[idx[j]] > sv[idx[i]]) { int t = idx[i]; idx[i] = idx[j]; idx[j] = t; }
  double Vs[9];
  for (int c = 0; c < 3; ++c) {
    S[c] = sv[idx[c]];
    for (int r = 0; r < 3; ++r) Vs[3 * r + c] = V[3 * r + idx[c]];
  }
  memcpy(V, Vs, sizeof(Vs));
}

#define R float
#define SUF f32
#include "oracle_impl.h"
#undef R
#undef SUF
#define R double
#define SUF f64
#include "oracle_impl.h"
#undef R
#undef SUF

/* mu = E / (2 (1 + nu)); lambda = 2 mu / (1 - 2 nu)  (IsotropicNeohookeanMaterial.cpp:15-16,
 * CUDAInterface.cu:704-705; the textbook lambda has an extra factor nu). */
void orc_lame(float E, float nu, float *mu, float *lambda) {
  float m = (float)(0.5 * (double)E / (double)(1 + nu));
  *mu = m;
  *lambda = m * 2 / (1 - 2 * nu);
}

/* Rest-shape precompute (TetMesh.cpp:189-216, computeANs :279-294):
 * Dm = [x1-x0, x2-x0, x3-x0]; vol = |det Dm|/6; m_a += rho vol/4 (in tet order);
 * Dm^-1; AN column i = -(v1 x v2 + v2 x v3 + v3 x v1)/6 over the edge table
 * {{1,0,2,3},{2,0,3,1},{3,0,1,2}}; gravity force = mass * grav taken before fixed nodes get
 * the 1e8 mass (:205, :212-216). */
void orc_precompute(int nn, const float *rest, int nt, const int *tets, const int *mask, float density,
                    const float grav[3], float fixed_mass, float *dminv, float *vol, float *an, float *mass,
                    float *gforce) {
  for (int i = 0; i < nn; ++i) mass[i] = 0;
  static const int edge[3][4] = {{1, 0, 2, 3}, {2, 0, 3, 1}, {3, 0, 1, 2}};
  for (int t = 0; t < nt; ++t) {
    const int *q = tets + 4 * t;
    float Dm[9];
    for (int k = 0; k < 3; ++k) {
      Dm[3 * k + 0] = rest[3 * q[1] + k] - rest[3 * q[0] + k];
      Dm[3 * k + 1] = rest[3 * q[2] + k] - rest[3 * q[0] + k];
      Dm[3 * k + 2] = rest[3 * q[3] + k] - rest[3 * q[0] + k];
    }
    float v = fabsf(det3_f32(Dm)) / 6.0f;
    vol[t] = v;
    float w = v / 4.0f * density;
    for (int j = 0; j < 4; ++j) mass[q[j]] += w;
    inv3_f32(Dm, dminv + 9 * t);
    for (int i = 0; i < 3; ++i) {
      const float *p0 = rest + 3 * q[edge[i][0]];
      float v1[3], v2[3], v3[3], c[3];
      for (int k = 0; k < 3; ++k) {
        v1[k] = rest[3 * q[edge[i][1]] + k] - p0[k];
        v2[k] = rest[3 * q[edge[i][2]] + k] - p0[k];
        v3[k] = rest[3 * q[edge[i][3]] + k] - p0[k];
      }
      c[0] = (v1[1] * v2[2] - v1[2] * v2[1]) + (v2[1] * v3[2] - v2[2] * v3[1]) + (v3[1] * v1[2] - v3[2] * v1[1]);
      c[1] = (v1[2] * v2[0] - v1[0] * v2[2]) + (v2[2] * v3[0] - v2[0] * v3[2]) + (v3[2] * v1[0] - v3[0] * v1[2]);
      c[2] = (v1[0] * v2[1] - v1[1] * v2[0]) + (v2[0] * v3[1] - v2[1] * v3[0]) + (v3[0] * v1[1] - v3[1] * v1[0]);
      for (int k = 0; k < 3; ++k) an[9 * t + 3 * k + i] = -c[k] / 6.0f;
    }
  }
  for (int i = 0; i < nn; ++i) {
    for (int k = 0; k < 3; ++k) gforce[3 * i + k] = grav[k] * mass[i];
    if (mask[i]) mass[i] = fixed_mass;
  }
}

/* Stiffness-matrix pattern (IsotropicMaterial.cpp:585-613): triplets (3a+i, 3b+j) for all node
 * pairs (a,b) of every tet, compressed row-major with sorted, de-duplicated columns
 * (Eigen::SparseMatrix<RowMajor>::setFromTriplets + makeCompressed).  diag_idx and kid use
 * the reference's binary search getCSRvalueIndex (IsotropicMaterial.h:120-138). */
static int csr_value_index(int i, int j, const int *rowptr, const int *col) {
  int l = rowptr[i], h = rowptr[i + 1] - 1, k = (l + h) / 2;
  while (l < h) {
    int kc = col[k];
    if (kc == j) break;
    else if (kc < j) l = k + 1;
    else h = k - 1;
    k = (l + h) / 2;
  }
  return k;
}

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

long orc_csr_pattern(int nn, int nt, const int *tets, int *rowptr, int *colidx, int *diag_idx, int *kid) {
  /* node adjacency (including self), sorted unique */
  int *deg = (int *)calloc((size_t)nn + 1, sizeof(int));
  for (int t = 0; t < nt; ++t)
    for (int a = 0; a < 4; ++a) deg[tets[4 * t + a]] += 4;
  int *off = (int *)malloc(sizeof(int) * ((size_t)nn + 1));
  off[0] = 0;
  for (int i = 0; i < nn; ++i) off[i + 1] = off[i] + deg[i];
  int *nb = (int *)malloc(sizeof(int) * (size_t)(off[nn] > 0 ? off[nn] : 1));
  int *fill = (int *)calloc((size_t)nn, sizeof(int));
  for (int t = 0; t < nt; ++t)
    for (int a = 0; a < 4; ++a) {
      int na = tets[4 * t + a];
      for (int b = 0; b < 4; ++b) nb[off[na] + fill[na]++] = tets[4 * t + b];
    }
  int *ucount = (int *)calloc((size_t)nn, sizeof(int));
  for (int i = 0; i < nn; ++i) {
    int *s = nb + off[i];
    int c = off[i + 1] - off[i];
    qsort(s, (size_t)c, sizeof(int), cmp_int);
    int u = 0;
    for (int j = 0; j < c; ++j)
      if (j == 0 || s[j] != s[j - 1]) s[u++] = s[j];
    ucount[i] = u;
  }
  long nnz = 0;
  for (int i = 0; i < nn; ++i) nnz += 9L * ucount[i];
  if (rowptr && colidx) {
    long pos = 0;
    for (int i = 0; i < nn; ++i)
      for (int d = 0; d < 3; ++d) {
        rowptr[3 * i + d] = (int)pos;
        for (int j = 0; j < ucount[i]; ++j)
          for (int e = 0; e < 3; ++e) colidx[pos++] = 3 * nb[off[i] + j] + e;
      }
    rowptr[3 * nn] = (int)pos;
    if (diag_idx)
      for (int r = 0; r < 3 * nn; ++r) diag_idx[r] = csr_value_index(r, r, rowptr, colidx);
    if (kid)
      for (int t = 0; t < nt; ++t)
        for (int fi = 0; fi < 4; ++fi)
          for (int fj = 0; fj < 3; ++fj) {
            int gki = tets[4 * t + fi] * 3 + fj, ki = fi * 3 + fj;
            for (int ni = 0; ni < 4; ++ni)
              kid[48 * t + ki * 4 + ni] = csr_value_index(gki, tets[4 * t + ni] * 3, rowptr, colidx);
          }
  }
  free(deg); free(off); free(nb); free(fill); free(ucount);
  return nnz;
}

/* Per-node incidence (ProjDynamic.cpp:36-42): entry 12*tet + 3*local appended in tet order. */
void orc_incidence(int nn, int nt, const int *tets, int *offsets, int *entries) {
  for (int i = 0; i <= nn; ++i) offsets[i] = 0;
  for (int t = 0; t < nt; ++t)
    for (int j = 0; j < 4; ++j) offsets[tets[4 * t + j] + 1]++;
  for (int i = 0; i < nn; ++i) offsets[i + 1] += offsets[i];
  int *fill = (int *)calloc((size_t)nn + 1, sizeof(int));
  for (int t = 0; t < nt; ++t)
    for (int j = 0; j < 4; ++j) {
      int n = tets[4 * t + j];
      entries[offsets[n] + fill[n]++] = 12 * t + 3 * j;
    }
  free(fill);
}

void orc_default_params(orc_params *p) {
  memset(p, 0, sizeof(*p));
  p->dt = 0.03f;
  p->grav[0] = 0; p->grav[1] = -9.8f; p->grav[2] = 0;
  p->alpha0 = 0.1f;
  p->alpha_decay = 0.7f;
  p->alpha_min = 0.01f;
  p->hd_refresh = 32;
  p->energy_check = 8;
  p->damping = 0.99f;
  p->init_momentum = 0.2f;
  p->profile_n = 3;
  p->profile_k[0] = 1; p->profile_k[1] = 7; p->profile_k[2] = 10;
  p->profile_rho[0] = 0.96f; p->profile_rho[1] = 0.991f; p->profile_rho[2] = 0.999f;
  p->semantics = ORC_REF_GPU;
  p->iterations = 96;
  p->relgrad_tol = 0;
  p->trace_cap = 0;
}

orc_solver *orc_create(int nn, const float *rest, int nt, const int *tets, const int *mask, float E, float nu,
                       float density, float fixed_mass, int material, int precision, const orc_params *p) {
  orc_solver *s = (orc_solver *)calloc(1, sizeof(orc_solver));
  s->nn = nn;
  s->nt = nt;
  s->material = material;
  s->precision = precision;
  s->p = *p;
  size_t T = (size_t)(nt > 0 ? nt : 1);
  s->tets = (int *)malloc(sizeof(int) * 4 * T);
  memcpy(s->tets, tets, sizeof(int) * 4 * (size_t)nt);
  s->mask = (int *)malloc(sizeof(int) * (size_t)nn);
  memcpy(s->mask, mask, sizeof(int) * (size_t)nn);
  s->dminv = (float *)malloc(sizeof(float) * 9 * T);
  s->an = (float *)malloc(sizeof(float) * 9 * T);
  s->vol = (float *)malloc(sizeof(float) * T);
  s->mass = (float *)malloc(sizeof(float) * (size_t)nn);
  s->gforce = (float *)malloc(sizeof(float) * 3 * (size_t)nn);
  orc_lame(E, nu, &s->mu, &s->lam);
  orc_precompute(nn, rest, nt, tets, mask, density, p->grav, fixed_mass, s->dminv, s->vol, s->an, s->mass,
                 s->gforce);
  s->inc_off = (int *)malloc(sizeof(int) * ((size_t)nn + 1));
  s->inc = (int *)malloc(sizeof(int) * 4 * T);
  orc_incidence(nn, nt, tets, s->inc_off, s->inc);
  s->omega = 1;
  if (precision == 64) {
    orc_state_f64 *st = (orc_state_f64 *)calloc(1, sizeof(orc_state_f64));
    state_alloc_f64(st, nn, nt);
    for (int i = 0; i < 3 * nn; ++i) st->x[i] = st->xlast[i] = rest[i];
    s->state = st;
  } else {
    orc_state_f32 *st = (orc_state_f32 *)calloc(1, sizeof(orc_state_f32));
    state_alloc_f32(st, nn, nt);
    for (int i = 0; i < 3 * nn; ++i) st->x[i] = st->xlast[i] = rest[i];
    s->state = st;
  }
  return s;
}

void orc_destroy(orc_solver *s) {
  if (!s) return;
  if (s->precision == 64) state_free_f64((orc_state_f64 *)s->state);
  else state_free_f32((orc_state_f32 *)s->state);
  free(s->state);
  free(s->tets); free(s->mask); free(s->dminv); free(s->an); free(s->vol); free(s->mass); free(s->gforce);
  free(s->inc_off); free(s->inc);
  free(s);
}

void orc_set_params(orc_solver *s, const orc_params *p) {
  float g[3] = {s->p.grav[0], s->p.grav[1], s->p.grav[2]};
  s->p = *p;
  /* gravity is baked into gforce; rescale when it changes */
  for (int i = 0; i < s->nn; ++i) {
    float m = 0;
    int ax = -1;
    for (int k = 0; k < 3; ++k)
      if (g[k] != 0) { ax = k; break; }
    if (ax >= 0) m = s->gforce[3 * i + ax] / g[ax];
    for (int k = 0; k < 3; ++k) s->gforce[3 * i + k] = p->grav[k] * m;
  }
}

void orc_set_threads(orc_solver *s, int nthreads) {
  (void)s;
#ifdef _OPENMP
  omp_set_num_threads(nthreads > 0 ? nthreads : 1);
#else
  (void)nthreads;
#endif
}

#define STATE_COPY_IN(T)                                                               \
  do {                                                                                 \
    T *st = (T *)s->state;                                                             \
    for (int i = 0; i < 3 * s->nn; ++i) {                                              \
      if (x) st->x[i] = x[i];                                                          \
      if (v) st->v[i] = v[i];                                                          \
      if (vl) st->vl[i] = vl[i];                                                       \
    }                                                                                  \
  } while (0)
#define STATE_COPY_OUT(T)                                                              \
  do {                                                                                 \
    T *st = (T *)s->state;                                                             \
    for (int i = 0; i < 3 * s->nn; ++i) {                                              \
      if (x) x[i] = st->x[i];                                                          \
      if (v) v[i] = st->v[i];                                                          \
      if (vl) vl[i] = st->vl[i];                                                       \
    }                                                                                  \
  } while (0)

void orc_set_state(orc_solver *s, const double *x, const double *v, const double *vl) {
  if (s->precision == 64) STATE_COPY_IN(orc_state_f64);
  else STATE_COPY_IN(orc_state_f32);
}

void orc_get_state(orc_solver *s, double *x, double *v, double *vl) {
  if (s->precision == 64) STATE_COPY_OUT(orc_state_f64);
  else STATE_COPY_OUT(orc_state_f32);
}

int orc_step(orc_solver *s, const double *ext, const double *x0_override, orc_stats *st, orc_trace *trace) {
  memset(st, 0, sizeof(*st));
  if (s->precision == 64) return step_f64(s, ext, x0_override, st, trace);
  return step_f32(s, ext, x0_override, st, trace);
}

double orc_gradient_norm(orc_solver *s, const double *x) {
  if (s->precision == 64) return grad_norm_at_f64(s, x);
  return grad_norm_at_f32(s, x);
}
