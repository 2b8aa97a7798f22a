/*
 * oracle_check.c -- drives every entry point of the CPU oracle (oracle/pbs_oracle.c) on a small
 * mesh so that the oracle can be built and run under AddressSanitizer / UndefinedBehaviorSanitizer
 * (tests/test_host_sanitized.py).  Test infrastructure only.
 *
 *   usage: oracle_check <mesh file>
 *   mesh file (native byte order): int32 nn, nt; float rest[3 nn]; int32 tets[4 nt];
 *   int32 mask[nn]; float x0[3 nn]
 * Prints a few sums; exit status 0 when every call returned normally with finite results.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pbs_oracle.h"

static int fail = 0;
static void finite(const char *what, const double *v, long n) {
  for (long i = 0; i < n; ++i)
    if (!isfinite(v[i])) {
      fprintf(stderr, "FAIL %s: non-finite value at %ld\n", what, i);
      fail = 1;
      return;
    }
}
static void finitef(const char *what, const float *v, long n) {
  for (long i = 0; i < n; ++i)
    if (!isfinite(v[i])) {
      fprintf(stderr, "FAIL %s: non-finite value at %ld\n", what, i);
      fail = 1;
      return;
    }
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  int nn = 0, nt = 0;
  if (fread(&nn, 4, 1, f) != 1 || fread(&nt, 4, 1, f) != 1 || nn <= 0 || nt <= 0) return 2;
  float *rest = malloc(sizeof(float) * 3 * nn), *x0 = malloc(sizeof(float) * 3 * nn);
  int *tets = malloc(sizeof(int) * 4 * nt), *mask = malloc(sizeof(int) * nn);
  if (fread(rest, 4, 3 * (size_t)nn, f) != 3 * (size_t)nn || fread(tets, 4, 4 * (size_t)nt, f) != 4 * (size_t)nt ||
      fread(mask, 4, (size_t)nn, f) != (size_t)nn || fread(x0, 4, 3 * (size_t)nn, f) != 3 * (size_t)nn)
    return 2;
  fclose(f);

  float mu, lam;
  orc_lame(1e5f, 0.45f, &mu, &lam);
  const float grav[3] = {0.0f, -9.8f, 0.0f};
  float *dminv = malloc(sizeof(float) * 9 * nt), *vol = malloc(sizeof(float) * nt), *an = malloc(sizeof(float) * 9 * nt);
  float *mass = malloc(sizeof(float) * nn), *gforce = malloc(sizeof(float) * 3 * nn);
  orc_precompute(nn, rest, nt, tets, mask, 1000.0f, grav, 1e8f, dminv, vol, an, mass, gforce);
  finitef("precompute", dminv, 9L * nt);

  double *x64 = malloc(sizeof(double) * 3 * nn), *v64 = calloc(3 * (size_t)nn, sizeof(double));
  for (long i = 0; i < 3L * nn; ++i) x64[i] = x0[i];
  float *tf32 = malloc(sizeof(float) * 12 * nt), *te32 = malloc(sizeof(float) * nt);
  double *tf64 = malloc(sizeof(double) * 12 * nt), *te64 = malloc(sizeof(double) * nt);
  float *hd32 = malloc(sizeof(float) * 3 * nn), *th32 = malloc(sizeof(float) * 12 * nt);
  double *hd64 = malloc(sizeof(double) * 3 * nn), *th64 = malloc(sizeof(double) * 12 * nt);
  for (int mat = 0; mat < 2; ++mat) {
    for (int clamp = 0; clamp < 3; ++clamp) {
      orc_tet_forces_f32(nt, tets, x0, dminv, an, vol, mu, lam, mat, clamp, tf32, te32);
      orc_tet_forces_f64(nt, tets, x64, dminv, an, vol, mu, lam, mat, clamp, tf64, te64);
    }
    finitef("tet_forces_f32", tf32, 12L * nt);
    finite("tet_forces_f64", tf64, 12L * nt);
    orc_hessian_diag_f32(nn, nt, tets, x0, dminv, an, mu, lam, mat, hd32, th32);
    orc_hessian_diag_f64(nn, nt, tets, x64, dminv, an, mu, lam, mat, hd64, th64);
    finite("hessian_diag_f64", hd64, 3L * nn);
  }
  /* SVD, dP/dF and the PD rotation on a few deformation gradients, including F = I (repeated
   * singular values: the perturbation branch) and an inverted one */
  {
    const double Fs[3][9] = {{1, 0, 0, 0, 1, 0, 0, 0, 1}, {1.1, 0.2, 0, -0.1, 0.9, 0.05, 0.02, 0, 1.3},
                             {-1, 0, 0, 0, 1, 0, 0, 0, 1}};
    for (int k = 0; k < 3; ++k) {
      double U[9], S[3], V[9], dp[81], Rt[9];
      float Ff[9], Uf[9], Sf[3], Vf[9], dpf[81], Rf[9];
      for (int i = 0; i < 9; ++i) Ff[i] = (float)Fs[k][i];
      orc_svd3_f64(Fs[k], U, S, V);
      orc_svd3_f32(Ff, Uf, Sf, Vf);
      orc_dpdf_f64(Fs[k], mu, lam, 0, dp);
      orc_dpdf_f32(Ff, mu, lam, 0, dpf);
      orc_pd_rotation_f64(Fs[k], Rt);
      orc_pd_rotation_f32(Ff, Rf);
      finite("svd3_f64", S, 3);
      finite("pd_rotation_f64", Rt, 9);
    }
  }
  /* adjacency, backward-Euler system, projective dynamics */
  const long nnz = orc_csr_pattern(nn, nt, tets, NULL, NULL, NULL, NULL);
  int *rowptr = malloc(sizeof(int) * (3 * (size_t)nn + 1)), *col = malloc(sizeof(int) * nnz);
  int *diag = malloc(sizeof(int) * 3 * (size_t)nn), *kid = malloc(sizeof(int) * 48 * (size_t)nt);
  orc_csr_pattern(nn, nt, tets, rowptr, col, diag, kid);
  int *ioff = malloc(sizeof(int) * (nn + 1)), *ient = malloc(sizeof(int) * 4 * (size_t)nt);
  orc_incidence(nn, nt, tets, ioff, ient);
  double *kval = malloc(sizeof(double) * nnz), *aval = malloc(sizeof(double) * nnz), *b = malloc(sizeof(double) * 3 * nn);
  float *v32 = calloc(3 * (size_t)nn, sizeof(float));
  orc_implicit_system_f32(nn, rest, nt, tets, mask, dminv, an, mass, gforce, mu, lam, 0, x0, v32, NULL, 0.03f, 0.02f,
                          0.02f, nnz, kid, diag, kval, aval, b);
  orc_implicit_system_f64(nn, rest, nt, tets, mask, dminv, an, mass, gforce, mu, lam, 1, x64, v64, NULL, 0.03f, 0.02f,
                          0.02f, nnz, kid, diag, kval, aval, b);
  finite("implicit_system_f64", aval, nnz);
  double *proj = malloc(sizeof(double) * 3 * nn);
  float *proj32 = malloc(sizeof(float) * 3 * nn);
  orc_pd_project_f64(nn, nt, tets, x64, dminv, vol, 1e6f, NULL, proj);
  orc_pd_project_f32(nn, nt, tets, x0, dminv, vol, 1e6f, NULL, proj32);
  orc_pd_matrix_f64(nn, nt, tets, dminv, vol, mass, 1e6f, 0.03f, nnz, kid, diag, kval);
  orc_pd_matrix_f32(nn, nt, tets, dminv, vol, mass, 1e6f, 0.03f, nnz, kid, diag, kval);
  finite("pd_project_f64", proj, 3L * nn);

  /* descent solver: both precisions, both semantics, both materials, fixed and rel-grad modes,
   * a caller's initial guess, external forces, the pass trace */
  double sum = 0.0;
  double *ext = calloc(3 * (size_t)nn, sizeof(double));
  ext[3 * (nn - 1) + 1] = 0.5;
  for (int prec = 32; prec <= 64; prec += 32)
    for (int sem = 0; sem < 2; ++sem)
      for (int mat = 0; mat < 2; ++mat) {
        orc_params p;
        orc_default_params(&p);
        p.semantics = sem;
        p.iterations = 40;
        p.trace_cap = 64;
        orc_trace tr[64];
        orc_solver *s = orc_create(nn, rest, nt, tets, mask, 1e5f, 0.45f, 1000.0f, 1e8f, mat, prec, &p);
        orc_set_threads(s, 2);
        orc_set_state(s, x64, NULL, NULL);
        orc_stats st;
        orc_step(s, NULL, NULL, &st, tr);
        orc_step(s, ext, NULL, &st, tr);
        p.relgrad_tol = 1e-3;
        p.iterations = 3000;
        orc_set_params(s, &p);
        orc_step(s, NULL, x64, &st, NULL);
        sum += orc_gradient_norm(s, x64);
        double *xs = malloc(sizeof(double) * 3 * nn), *vs = malloc(sizeof(double) * 3 * nn),
               *vl = malloc(sizeof(double) * 3 * nn);
        orc_get_state(s, xs, vs, vl);
        finite("solver state", xs, 3L * nn);
        for (long i = 0; i < 3L * nn; ++i) sum += xs[i];
        free(xs);
        free(vs);
        free(vl);
        orc_destroy(s);
      }
  printf("oracle_check: nn %d nt %d nnz %ld state sum %.6e\n", nn, nt, nnz, sum);
  free(rest); free(x0); free(tets); free(mask); free(dminv); free(vol); free(an); free(mass); free(gforce);
  free(x64); free(v64); free(tf32); free(te32); free(tf64); free(te64); free(hd32); free(th32); free(hd64);
  free(th64); free(rowptr); free(col); free(diag); free(kid); free(ioff); free(ient); free(kval); free(aval);
  free(b); free(v32); free(proj); free(proj32); free(ext);
  if (fail) return 1;
  printf("oracle_check: all passed\n");
  return 0;
}
