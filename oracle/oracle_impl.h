/*
 * oracle_impl.h -- precision-generic body of the PhyBS descent oracle.
 * TEST INFRASTRUCTURE ONLY (see pbs_oracle.h).  Included twice by pbs_oracle.c,
 * with R = float / SUF = f32 and R = double / SUF = f64.
 */
#define ORC_CAT_(a, b) a##_##b
#define ORC_CAT(a, b) ORC_CAT_(a, b)
#define FN(name) ORC_CAT(name, SUF)

/* cuMath::Matrix_Determinant_3 (CUDA_MATH.h:157-161): cofactor expansion down column 0. */
static inline R FN(det3)(const R *x) {
  return x[0] * (x[4] * x[8] - x[7] * x[5]) + x[3] * (x[7] * x[2] - x[1] * x[8]) +
         x[6] * (x[1] * x[5] - x[4] * x[2]);
}

/* cuMath::Matrix_Inverse_3 (CUDA_MATH.h:140-155): adjugate / determinant, row-major. */
static inline void FN(inv3)(const R *A, R *Rm) {
  Rm[0] = A[4] * A[8] - A[7] * A[5];
  Rm[1] = A[7] * A[2] - A[1] * A[8];
  Rm[2] = A[1] * A[5] - A[4] * A[2];
  Rm[3] = A[5] * A[6] - A[3] * A[8];
  Rm[4] = A[0] * A[8] - A[2] * A[6];
  Rm[5] = A[2] * A[3] - A[0] * A[5];
  Rm[6] = A[3] * A[7] - A[4] * A[6];
  Rm[7] = A[1] * A[6] - A[0] * A[7];
  Rm[8] = A[0] * A[4] - A[1] * A[3];
  R det = A[0] * Rm[0] + A[3] * Rm[1] + A[6] * Rm[2];
  R inv = (R)1 / det;
  for (int i = 0; i < 9; ++i) Rm[i] *= inv;
}

static inline void FN(mul3)(const R *A, const R *B, R *C) { /* C = A B, row-major */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

/* Deformation gradient F = Ds Dm^-1 with Ds columns x_{1,2,3} - x_0
 * (TetMesh.cpp:307-317; CUDAInterface.cu:506-516). */
static inline void FN(def_grad)(const int *t, const R *x, const float *dminv, R *F) {
  R Ds[9], Dmi[9];
  for (int k = 0; k < 3; ++k) {
    Ds[3 * k + 0] = x[3 * t[1] + k] - x[3 * t[0] + k];
    Ds[3 * k + 1] = x[3 * t[2] + k] - x[3 * t[0] + k];
    Ds[3 * k + 2] = x[3 * t[3] + k] - x[3 * t[0] + k];
  }
  for (int i = 0; i < 9; ++i) Dmi[i] = (R)dminv[i];
  FN(mul3)(Ds, Dmi, F);
}

/* 3x3 SVD with the reference's conventions (IsotropicMaterial.cpp:16-60):
 * singular values descending, det V = +1 (negate V column 0), U = F V S^-1 with 1e-8
 * guards, degenerate U completed by cross products, det U = +1 (negate U column 2 and S2).
 * The unconstrained decomposition (the reference calls Eigen::JacobiSVD) is computed by
 * one-sided Jacobi in FP64 (svd3_core in pbs_oracle.c) and rounded to R. */
void FN(orc_svd3)(const R *F, R *U, R *S, R *V) {
  double Fd[9], Sd[3], Vd[9];
  for (int i = 0; i < 9; ++i) Fd[i] = (double)F[i];
  svd3_core(Fd, Sd, Vd);
  for (int i = 0; i < 3; ++i) S[i] = (R)Sd[i];
  for (int i = 0; i < 9; ++i) V[i] = (R)Vd[i];
  if (FN(det3)(V) < 0) { V[0] = -V[0]; V[3] = -V[3]; V[6] = -V[6]; }
  const R eps = (R)1e-8; /* m_eps_singularvalue (IsotropicMaterial.cpp:4) */
  R ss[3];
  for (int i = 0; i < 3; ++i) ss[i] = S[i] > eps ? (R)1 / S[i] : (R)0;
  R FV[9];
  FN(mul3)(F, V, FV);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) U[3 * r + c] = FV[3 * r + c] * ss[c];
  if (S[0] < eps) {
    for (int i = 0; i < 9; ++i) U[i] = (i % 4 == 0) ? (R)1 : (R)0;
  } else if (S[1] < eps) {
    R u0[3] = {U[0], U[3], U[6]}, v1[3], v2[3];
    double o[3];
    unit_orthogonal((double)u0[0], (double)u0[1], (double)u0[2], o);
    for (int r = 0; r < 3; ++r) v1[r] = (R)o[r];
    v2[0] = u0[1] * v1[2] - u0[2] * v1[1];
    v2[1] = u0[2] * v1[0] - u0[0] * v1[2];
    v2[2] = u0[0] * v1[1] - u0[1] * v1[0];
    R n = (R)sqrt((double)(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]));
    for (int r = 0; r < 3; ++r) { U[3 * r + 1] = v1[r]; U[3 * r + 2] = v2[r] / n; }
  } else if (S[2] < eps) {
    R c[3];
    c[0] = U[3] * U[7] - U[6] * U[4];
    c[1] = U[6] * U[1] - U[0] * U[7];
    c[2] = U[0] * U[4] - U[3] * U[1];
    R n = (R)sqrt((double)(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]));
    for (int r = 0; r < 3; ++r) U[3 * r + 2] = c[r] / n;
  }
  if (FN(det3)(U) < 0) { U[2] = -U[2]; U[5] = -U[5]; U[8] = -U[8]; S[2] = -S[2]; }
}

/* Principal-stretch energy gradient and Hessian.
 * Neo-Hookean: IsotropicNeohookeanMaterial.cpp:50-91 (CUDAInterface.cu:10-46).
 * Stretch material psi = sum mu (s_i-1)^2 + lambda/2 (J-1)^2: no reference counterpart;
 * derivatives follow from the formula (dpsi/ds_i = 2mu(s_i-1) + lambda (J-1) J/s_i). */
static void FN(fhat_grad_hess)(const R *s, R mu, R lam, int material, R *g, R *h) {
  if (material == ORC_MAT_NEOHOOKEAN) {
    R lsum = (R)log((double)s[0]) + (R)log((double)s[1]) + (R)log((double)s[2]);
    R tmp = lam * lsum - mu;
    for (int i = 0; i < 3; ++i) g[i] = mu * s[i] + tmp / s[i];
    R t = lam * lsum;
    R t1 = mu + lam - t, t2 = t - mu;
    R i0 = (R)1 / s[0], i1 = (R)1 / s[1], i2 = (R)1 / s[2];
    h[0] = mu + t1 * i0 * i0;
    h[4] = mu + t1 * i1 * i1;
    h[8] = mu + t1 * i2 * i2;
    h[1] = h[3] = (t1 + t2) * i0 * i1;
    h[2] = h[6] = (t1 + t2) * i0 * i2;
    h[5] = h[7] = (t1 + t2) * i1 * i2;
  } else {
    R J = s[0] * s[1] * s[2];
    for (int i = 0; i < 3; ++i) g[i] = (R)2 * mu * (s[i] - (R)1) + lam * (J - (R)1) * J / s[i];
    for (int i = 0; i < 3; ++i) h[4 * i] = (R)2 * mu + lam * (J / s[i]) * (J / s[i]);
    R c = lam * J * ((R)2 * J - (R)1);
    h[1] = h[3] = c / (s[0] * s[1]);
    h[2] = h[6] = c / (s[0] * s[2]);
    h[5] = h[7] = c / (s[1] * s[2]);
  }
}

/* dP/dF_ij by Xu et al. 2015 eq. 5 (IsotropicMaterial.cpp:963-1020 and :856-870).
 * U, V row-major; the 2x2 systems are solved in FP64 as the CPU reference does
 * (Eigen::Matrix2d inverse, IsotropicMaterial.cpp:977-992). */
static void FN(dpdf_ij)(const R *U, const R *s, const R *V, const R *Ph, const R *H, int i, int j, R *out) {
  R wU[9] = {0}, wVT[9] = {0};
  static const int kl[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int q = 0; q < 3; ++q) {
    int k = kl[q][0], l = kl[q][1];
    double a = (double)s[l], b = (double)s[k];
    double u0 = (double)(U[3 * i + k] * V[3 * j + l]);
    double u1 = (double)(-U[3 * i + l] * V[3 * j + k]);
    double det = a * a - b * b; /* inverse of [[a,b],[b,a]] = [[a,-b],[-b,a]]/det */
    double w0 = (a * u0 - b * u1) / det;
    double w1 = (-b * u0 + a * u1) / det;
    wU[3 * k + l] = (R)w0;
    wU[3 * l + k] = -(R)w0;
    wVT[3 * k + l] = (R)w1;
    wVT[3 * l + k] = -(R)w1;
  }
  R dU[9], dVT[9], VT[9];
  FN(mul3)(U, wU, dU);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) VT[3 * r + c] = V[3 * c + r];
  FN(mul3)(wVT, VT, dVT);
  R w[3], dPh[3];
  for (int k = 0; k < 3; ++k) w[k] = U[3 * i + k] * V[3 * j + k];
  for (int k = 0; k < 3; ++k) dPh[k] = H[k] * w[0] + H[k + 3] * w[1] + H[k + 6] * w[2];
  /* dU diag(Ph) V^T + U diag(dPh) V^T + U diag(Ph) dVT */
  R A[9], B[9], C[9], T1[9], T2[9], T3[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      A[3 * r + c] = dU[3 * r + c] * Ph[c];
      B[3 * r + c] = U[3 * r + c] * dPh[c];
      C[3 * r + c] = U[3 * r + c] * Ph[c];
    }
  FN(mul3)(A, VT, T1);
  FN(mul3)(B, VT, T2);
  FN(mul3)(C, dVT, T3);
  for (int m = 0; m < 9; ++m) out[m] = T1[m] + T2[m] + T3[m];
}

/* Full dP/dF with the degenerate-singular-value perturbation (IsotropicMaterial.cpp:872-961;
 * CUDAInterface.cu:101-176).  out[k*9 + Fid] = dP_k / dF_Fid. */
void FN(orc_dpdf)(const R *F, float mu_f, float lam_f, int material, R *out) {
  R mu = (R)mu_f, lam = (R)lam_f;
  R U[9], s[3], V[9];
  FN(orc_svd3)(F, U, s, V);
  const R eps = (R)1e-6;
  R ft[3];
  int perturbed = 1;
  if (s[0] - s[2] < (R)2 * eps) {
    ft[2] = s[2];
    ft[1] = ft[2] + eps;
    ft[0] = ft[1] + eps;
  } else if ((s[0] - s[1] < eps) && (s[1] - s[2] >= eps)) {
    ft[2] = s[2];
    ft[1] = s[0] - eps;
    ft[0] = s[0];
  } else if ((s[0] - s[1] >= eps) && (s[1] - s[2] < eps)) {
    ft[2] = s[2];
    ft[1] = s[2] + eps;
    ft[0] = s[0];
  } else {
    ft[0] = s[0];
    ft[1] = s[1];
    ft[2] = s[2];
    perturbed = 0;
  }
  R Un[9], sn[3], Vn[9];
  if (perturbed) {
    R Fn[9], US[9], VT[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) { US[3 * r + c] = U[3 * r + c] * ft[c]; VT[3 * r + c] = V[3 * c + r]; }
    FN(mul3)(US, VT, Fn);
    FN(orc_svd3)(Fn, Un, sn, Vn);
  } else {
    memcpy(Un, U, sizeof(Un));
    memcpy(Vn, V, sizeof(Vn));
    memcpy(sn, s, sizeof(sn));
  }
  R Ph[3], H[9];
  FN(fhat_grad_hess)(sn, mu, lam, material, Ph, H);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      R d[9];
      FN(dpdf_ij)(Un, sn, Vn, Ph, H, i, j, d);
      for (int k = 0; k < 9; ++k) out[k * 9 + 3 * i + j] = d[k];
    }
}

/* Elastic force and energy of one tet.
 * Neo-Hookean (IsotropicNeohookeanMaterial.cpp:99-111, :122-129): P = mu (F - F^-T) + lambda ln J F^-T,
 * H = P AN, node1..3 += columns of H, node0 -= row sums; psi V with the chosen clamp.
 * The FP32 energy keeps the reference's double-promoted literals (0.5 * ...).
 * Stretch material: P = 2 mu (F - R) + lambda (J-1) J F^-T with R = U V^T from orc_svd3. */
static void FN(tet_force)(const R *F, const float *anf, float volf, R mu, R lam, int material, int clamp,
                          R *f12, R *energy) {
  R Fi[9], FiT[9], P[9], AN[9], G[9];
  FN(inv3)(F, Fi);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) FiT[3 * r + c] = Fi[3 * c + r];
  R J = FN(det3)(F);
  R e;
  if (material == ORC_MAT_NEOHOOKEAN) {
    R lj = (R)log((double)J);
    for (int i = 0; i < 9; ++i) P[i] = (F[i] - FiT[i]) * mu + lam * lj * FiT[i];
    R I1 = 0;
    for (int i = 0; i < 9; ++i) I1 += F[i] * F[i];
    e = (R)(0.5 * (double)mu * (double)(I1 - (R)3) - (double)(mu * lj) + 0.5 * (double)lam * (double)lj * (double)lj);
  } else {
    R U[9], s[3], V[9], Rot[9], VT[9];
    FN(orc_svd3)(F, U, s, V);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) VT[3 * r + c] = V[3 * c + r];
    FN(mul3)(U, VT, Rot);
    R c = lam * (J - (R)1) * J;
    for (int i = 0; i < 9; ++i) P[i] = (R)2 * mu * (F[i] - Rot[i]) + c * FiT[i];
    R acc = 0;
    for (int i = 0; i < 3; ++i) acc += (s[i] - (R)1) * (s[i] - (R)1);
    e = mu * acc + (R)0.5 * lam * (J - (R)1) * (J - (R)1);
  }
  for (int i = 0; i < 9; ++i) AN[i] = (R)anf[i];
  FN(mul3)(P, AN, G);
  for (int a = 0; a < 3; ++a) {
    f12[3 + a] = G[3 * a + 0];
    f12[6 + a] = G[3 * a + 1];
    f12[9 + a] = G[3 * a + 2];
    f12[a] = -(G[3 * a + 0] + G[3 * a + 1] + G[3 * a + 2]);
  }
  e *= (R)volf;
  if (clamp == ORC_CLAMP_GPU) e = e > (R)0 ? e : (R)0;
  else if (clamp == ORC_CLAMP_CPU) e = e > (R)1e-6 ? e : (R)0;
  *energy = e;
}

void FN(orc_tet_forces)(int nt, const int *tets, const R *x, const float *dminv, const float *an,
                        const float *vol, float mu, float lam, int material, int clamp, R *tetf, R *tete) {
#pragma omp parallel for schedule(static)
  for (int t = 0; t < nt; ++t) {
    R F[9], f12[12], e;
    FN(def_grad)(tets + 4 * t, x, dminv + 9 * t, F);
    FN(tet_force)(F, an + 9 * t, vol[t], (R)mu, (R)lam, material, clamp, f12, &e);
    if (tetf) memcpy(tetf + 12 * t, f12, sizeof(f12));
    if (tete) tete[t] = e;
  }
}

/* Per-tet stiffness diagonal K_kk = -df_k/dx_k for the 12 DOFs, following the product
 * chain dGdF = blkdiag(AN^T) dPdF, dGdx = dGdF dFdx, K = -dfdx of
 * IsotropicMaterial.cpp:503-552 restricted to its diagonal:
 *   K(n,d) = -sum_{c,b} an_n[c] dP_(d,b)/dF_(d,c) c_n[b]
 * with an_n / c_n the node-n columns of AN / rows of Dm^-1 (node 0: minus their sums). */
static void FN(tet_hd)(const R *F, const float *dmi, const float *an, float mu, float lam, int material, R *h12) {
  R M[81];
  FN(orc_dpdf)(F, mu, lam, material, M);
  for (int n = 0; n < 4; ++n) {
    R anv[3], cv[3];
    for (int c = 0; c < 3; ++c) {
      if (n == 0) {
        anv[c] = -((R)an[3 * c] + (R)an[3 * c + 1] + (R)an[3 * c + 2]);
        cv[c] = -((R)dmi[c] + (R)dmi[3 + c] + (R)dmi[6 + c]);
      } else {
        anv[c] = (R)an[3 * c + n - 1];
        cv[c] = (R)dmi[3 * (n - 1) + c];
      }
    }
    for (int d = 0; d < 3; ++d) {
      R acc = 0;
      for (int b = 0; b < 3; ++b) {
        R inner = 0;
        for (int c = 0; c < 3; ++c) inner += anv[c] * M[(3 * d + b) * 9 + 3 * d + c];
        acc += inner * cv[b];
      }
      h12[3 * n + d] = -acc;
    }
  }
}

void FN(orc_hessian_diag)(int nn, int nt, const int *tets, const R *x, const float *dminv, const float *an,
                          float mu, float lam, int material, R *hd, R *tethd) {
  R *per = (R *)malloc(sizeof(R) * 12 * (size_t)(nt > 0 ? nt : 1));
#pragma omp parallel for schedule(static)
  for (int t = 0; t < nt; ++t) {
    R F[9];
    FN(def_grad)(tets + 4 * t, x, dminv + 9 * t, F);
    FN(tet_hd)(F, dminv + 9 * t, an + 9 * t, mu, lam, material, per + 12 * t);
  }
  if (hd) {
    for (int i = 0; i < 3 * nn; ++i) hd[i] = 0;
    for (int t = 0; t < nt; ++t)
      for (int n = 0; n < 4; ++n)
        for (int d = 0; d < 3; ++d) hd[3 * tets[4 * t + n] + d] += per[12 * t + 3 * n + d];
  }
  if (tethd) memcpy(tethd, per, sizeof(R) * 12 * (size_t)nt);
  free(per);
}

/* Full per-tet stiffness K = -df/dx (12 x 12, row 3n+d, column 3m+e), by the reference's chain
 * (compute_dfdx, CUDAInterface.cu:178-219; IsotropicMaterial.cpp:503-552): with M = dP/dF from
 * orc_dpdf, an_n the AN column of node n and c_m the Dm^-1 row of node m (node 0: minus sums),
 *   K(n,d),(m,e) = -sum_{b,c} an_n[b] dP_(d,b)/dF_(e,c) c_m[c]. */
static void FN(tet_stiffness)(const R *F, const float *dmi, const float *an, float mu, float lam, int material,
                              R *K) {
  R M[81];
  FN(orc_dpdf)(F, mu, lam, material, M);
  R anv[4][3], cv[4][3];
  for (int n = 0; n < 4; ++n)
    for (int c = 0; c < 3; ++c) {
      if (n == 0) {
        anv[n][c] = -((R)an[3 * c] + (R)an[3 * c + 1] + (R)an[3 * c + 2]);
        cv[n][c] = -((R)dmi[c] + (R)dmi[3 + c] + (R)dmi[6 + c]);
      } else {
        anv[n][c] = (R)an[3 * c + n - 1];
        cv[n][c] = (R)dmi[3 * (n - 1) + c];
      }
    }
  for (int n = 0; n < 4; ++n)
    for (int d = 0; d < 3; ++d)
      for (int m = 0; m < 4; ++m)
        for (int e = 0; e < 3; ++e) {
          R acc = 0;
          for (int b = 0; b < 3; ++b) {
            R inner = 0;
            for (int c = 0; c < 3; ++c) inner += M[(3 * d + b) * 9 + 3 * e + c] * cv[m][c];
            acc += anv[n][b] * inner;
          }
          K[(3 * n + d) * 12 + 3 * m + e] = -acc;
        }
}

/* Backward-Euler linear system of the reference's GPU implicit path (fvm.cpp:179-183):
 *   computeElasticForces (CUDAInterface.cu:345-416): P = U diag(dpsi/ds) V^T from the SVD,
 *     forces = P AN scattered as in tet_force;
 *   addFixContraintsAndGravity (:327-343): fixed nodes f = 1e8 (rest - x), free nodes f += m g;
 *   computeGlobalK (:234-271): kval[kID[48t + 4(3 fi + fj) + ni] + nj] += K(fi,fj),(ni,nj);
 *   doBackEuler (:819-852): b = h f + m v (setRHSofLinearSystem), A = K (h + beta) h
 *     (scaleDeviceArray), A_ii += m (1 + alpha h) (setMassScaled + setLHSofLinearSystem).
 * ext (3N, may be NULL) is added to the free nodes' forces (the reference has no such input).
 * Sums run in ascending tet order in precision R; outputs are widened to double. */
void FN(orc_implicit_system)(int nn, const float *rest, int nt, const int *tets, const int *mask,
                             const float *dminv, const float *an, const float *mass, const float *gforce,
                             float mu_f, float lam_f, int material, const R *x, const R *v, const R *ext,
                             float h_f, float alpha_f, float beta_f, long nnz, const int *kid, const int *diag_idx,
                             double *kval, double *aval, double *b) {
  const R mu = (R)mu_f, lam = (R)lam_f, h = (R)h_f;
  R *kv = (R *)calloc((size_t)(nnz > 0 ? nnz : 1), sizeof(R));
  R *f = (R *)calloc(3 * (size_t)nn, sizeof(R));
  for (int t = 0; t < nt; ++t) {
    const int *q = tets + 4 * t;
    R F[9], U[9], s[3], V[9], ph[3], H[9], P[9], AN[9], G[9], K[144];
    FN(def_grad)(q, x, dminv + 9 * t, F);
    FN(orc_svd3)(F, U, s, V);
    FN(fhat_grad_hess)(s, mu, lam, material, ph, H);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c)
        P[3 * r + c] = U[3 * r] * ph[0] * V[3 * c] + U[3 * r + 1] * ph[1] * V[3 * c + 1] + U[3 * r + 2] * ph[2] * V[3 * c + 2];
    for (int i = 0; i < 9; ++i) AN[i] = (R)an[9 * t + i];
    FN(mul3)(P, AN, G);
    for (int a = 0; a < 3; ++a) {
      f[3 * q[1] + a] += G[3 * a + 0];
      f[3 * q[2] + a] += G[3 * a + 1];
      f[3 * q[3] + a] += G[3 * a + 2];
      f[3 * q[0] + a] += -(G[3 * a + 0] + G[3 * a + 1] + G[3 * a + 2]);
    }
    FN(tet_stiffness)(F, dminv + 9 * t, an + 9 * t, mu_f, lam_f, material, K);
    for (int fi = 0; fi < 4; ++fi)
      for (int fj = 0; fj < 3; ++fj)
        for (int ni = 0; ni < 4; ++ni)
          for (int nj = 0; nj < 3; ++nj)
            kv[kid[48 * (size_t)t + 4 * (3 * fi + fj) + ni] + nj] += K[(3 * fi + fj) * 12 + 3 * ni + nj];
  }
  const R s1 = h + (R)beta_f;
  for (long i = 0; i < nnz; ++i) {
    kval[i] = (double)kv[i];
    aval[i] = (double)(kv[i] * s1 * h);
  }
  const R ms = (R)1 + (R)alpha_f * h;
  for (int n = 0; n < nn; ++n)
    for (int d = 0; d < 3; ++d) {
      const long di = diag_idx[3 * n + d];
      aval[di] = (double)((R)aval[di] + (R)mass[n] * ms);
      R fk;
      if (mask[n]) {
        fk = (R)(1e8 * (double)((R)rest[3 * n + d] - x[3 * n + d]));
      } else {
        fk = f[3 * n + d] + (R)gforce[3 * n + d];
        if (ext) fk += ext[3 * n + d];
      }
      b[3 * n + d] = (double)(fk * h + (R)mass[n] * v[3 * n + d]);
    }
  free(kv);
  free(f);
}

/* Projection rotation of projective dynamics (ProjDynamic.cpp:359-368): the SVD F = U S V^T with
 * singular values in decreasing order (the reference calls Eigen::JacobiSVD; here the one-sided
 * Jacobi of svd3_core in FP64 with U = F V S^-1), then U's column 0 negated if det U < 0 and V's
 * column 0 negated if det V < 0, R = U V^T; rounded to R. */
void FN(orc_pd_rotation)(const R *F, R *Rot) {
  double Fd[9], S[3], V[9], U[9];
  for (int i = 0; i < 9; ++i) Fd[i] = (double)F[i];
  svd3_core(Fd, S, V);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double a = Fd[3 * r] * V[c] + Fd[3 * r + 1] * V[3 + c] + Fd[3 * r + 2] * V[6 + c];
      U[3 * r + c] = S[c] > 1e-300 ? a / S[c] : 0.0;
    }
  if (!(S[2] > 1e-300)) { /* complete a rank-2 U by the cross product of its first columns */
    U[2] = U[3] * U[7] - U[6] * U[4];
    U[5] = U[6] * U[1] - U[0] * U[7];
    U[8] = U[0] * U[4] - U[3] * U[1];
  }
  double du = U[0] * (U[4] * U[8] - U[7] * U[5]) + U[3] * (U[7] * U[2] - U[1] * U[8]) + U[6] * (U[1] * U[5] - U[4] * U[2]);
  double dv = V[0] * (V[4] * V[8] - V[7] * V[5]) + V[3] * (V[7] * V[2] - V[1] * V[8]) + V[6] * (V[1] * V[5] - V[4] * V[2]);
  if (du < 0) { U[0] = -U[0]; U[3] = -U[3]; U[6] = -U[6]; }
  if (dv < 0) { V[0] = -V[0]; V[3] = -V[3]; V[6] = -V[6]; }
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      Rot[3 * r + c] = (R)(U[3 * r] * V[3 * c] + U[3 * r + 1] * V[3 * c + 1] + U[3 * r + 2] * V[3 * c + 2]);
}

/* Local step of projective dynamics (projectLocalConstraints2, ProjDynamic.cpp:223-308): per tet
 * c = w Ac^T vec(R) with w = stiffness * vol, i.e. corner i gets w R c_i (c_0 = minus the column sums
 * of Dm^-1, c_i = row i-1); out = s + the corner projections summed per node in ascending tet
 * order, starting from s (s may be NULL: zero).  x, s, out: 3N AoS. */
void FN(orc_pd_project)(int nn, int nt, const int *tets, const R *x, const float *dminv, const float *vol,
                        float stiffness, const R *s, R *out) {
  R *proj = (R *)malloc(sizeof(R) * 12 * (size_t)(nt > 0 ? nt : 1));
#pragma omp parallel for schedule(static)
  for (int t = 0; t < nt; ++t) {
    R F[9], Rot[9];
    FN(def_grad)(tets + 4 * t, x, dminv + 9 * t, F);
    FN(orc_pd_rotation)(F, Rot);
    const float *D = dminv + 9 * t;
    R cv[4][3];
    for (int j = 0; j < 3; ++j) {
      cv[0][j] = -(R)D[j] - (R)D[3 + j] - (R)D[6 + j];
      cv[1][j] = (R)D[j];
      cv[2][j] = (R)D[3 + j];
      cv[3][j] = (R)D[6 + j];
    }
    const R w = (R)stiffness * (R)vol[t];
    for (int i = 0; i < 4; ++i)
      for (int r = 0; r < 3; ++r)
        proj[12 * (size_t)t + 3 * i + r] = (cv[i][0] * Rot[3 * r] + cv[i][1] * Rot[3 * r + 1] + cv[i][2] * Rot[3 * r + 2]) * w;
  }
  for (int i = 0; i < 3 * nn; ++i) out[i] = s ? s[i] : (R)0;
  for (int t = 0; t < nt; ++t)
    for (int i = 0; i < 4; ++i)
      for (int r = 0; r < 3; ++r) out[3 * tets[4 * t + i] + r] += proj[12 * (size_t)t + 3 * i + r];
  free(proj);
}

/* Global matrix of projective dynamics (buildGlobalSolverMatrix, ProjDynamic.cpp:46-96): the
 * explicit 9x12 Ac (rows 3a..3a+2 of column block i hold c_i), w Ac^T Ac added per tet via kid, and
 * m / t^2 on the diagonal; values on the CSR pattern of orc_csr_pattern. */
void FN(orc_pd_matrix)(int nn, int nt, const int *tets, const float *dminv, const float *vol, const float *mass,
                       float stiffness, float t_f, long nnz, const int *kid, const int *diag_idx, double *gval) {
  R *g = (R *)calloc((size_t)(nnz > 0 ? nnz : 1), sizeof(R));
  const R t = (R)t_f, t2inv = (R)1 / (t * t);
  for (int n = 0; n < nn; ++n)
    for (int d = 0; d < 3; ++d) g[diag_idx[3 * n + d]] += (R)mass[n] * t2inv;
  for (int k = 0; k < nt; ++k) {
    const float *D = dminv + 9 * k;
    R Ac[9][12];
    memset(Ac, 0, sizeof(Ac));
    for (int a = 0; a < 3; ++a)
      for (int r = 0; r < 3; ++r) {
        Ac[3 * a + r][a] = -(R)D[r] - (R)D[3 + r] - (R)D[6 + r];
        Ac[3 * a + r][3 + a] = (R)D[r];
        Ac[3 * a + r][6 + a] = (R)D[3 + r];
        Ac[3 * a + r][9 + a] = (R)D[6 + r];
      }
    const R w = (R)stiffness * (R)vol[k];
    for (int fi = 0; fi < 4; ++fi)
      for (int fj = 0; fj < 3; ++fj)
        for (int ni = 0; ni < 4; ++ni)
          for (int nj = 0; nj < 3; ++nj) {
            R acc = 0;
            for (int q = 0; q < 9; ++q) acc += Ac[q][3 * fi + fj] * Ac[q][3 * ni + nj];
            g[kid[48 * (size_t)k + 4 * (3 * fi + fj) + ni] + nj] += w * acc;
          }
  }
  for (long i = 0; i < nnz; ++i) gval[i] = (double)g[i];
  free(g);
}

/* ------------------------------------------------------------------------------------------
 * Descent solver state in precision R.
 * ------------------------------------------------------------------------------------------ */
typedef struct FN(orc_state) {
  R *x, *v, *vl, *xlast;          /* persistent: positions, velocities, v_last, Chebyshev x_last */
  R *xt, *xold, *xnext, *cur, *hd; /* per step */
  R *tetf, *tete, *tethd, *nodef, *g;
} FN(orc_state);

static void FN(state_alloc)(FN(orc_state) *st, int nn, int nt) {
  size_t n3 = 3 * (size_t)nn;
  st->x = (R *)calloc(n3, sizeof(R));
  st->v = (R *)calloc(n3, sizeof(R));
  st->vl = (R *)calloc(n3, sizeof(R));
  st->xlast = (R *)calloc(n3, sizeof(R));
  st->xt = (R *)calloc(n3, sizeof(R));
  st->xold = (R *)calloc(n3, sizeof(R));
  st->xnext = (R *)calloc(n3, sizeof(R));
  st->cur = (R *)calloc(n3, sizeof(R));
  st->hd = (R *)calloc(n3, sizeof(R));
  st->nodef = (R *)calloc(n3, sizeof(R));
  st->g = (R *)calloc(n3, sizeof(R));
  st->tetf = (R *)calloc(12 * (size_t)(nt > 0 ? nt : 1), sizeof(R));
  st->tete = (R *)calloc((size_t)(nt > 0 ? nt : 1), sizeof(R));
  st->tethd = (R *)calloc(12 * (size_t)(nt > 0 ? nt : 1), sizeof(R));
}

static void FN(state_free)(FN(orc_state) *st) {
  free(st->x); free(st->v); free(st->vl); free(st->xlast); free(st->xt); free(st->xold);
  free(st->xnext); free(st->cur); free(st->hd); free(st->nodef); free(st->g);
  free(st->tetf); free(st->tete); free(st->tethd);
}

/* Elastic forces at y: per-tet (tetf, tete) and gathered per node over the incidence list
 * in ascending tet order (ProjDynamic.cpp:25-45 ordering). */
static void FN(forces_at)(orc_solver *s, FN(orc_state) *st, const R *y) {
  FN(orc_tet_forces)(s->nt, s->tets, y, s->dminv, s->an, s->vol, s->mu, s->lam, s->material,
                     s->p.semantics == ORC_REF_CPU ? ORC_CLAMP_CPU : ORC_CLAMP_GPU, st->tetf, st->tete);
#pragma omp parallel for schedule(static)
  for (int n = 0; n < s->nn; ++n) {
    R acc[3] = {0, 0, 0};
    const int e0 = s->inc_off[n], e1 = s->inc_off[n + 1];
    for (int q = 0; q < e1 - e0; ++q) {
      int slot = s->inc[s->p.gather_order ? e1 - 1 - q : e0 + q]; /* 12*tet + 3*local */
      for (int d = 0; d < 3; ++d) acc[d] += st->tetf[slot + d];
    }
    for (int d = 0; d < 3; ++d) st->nodef[3 * n + d] = acc[d];
  }
}

static void FN(hd_at)(orc_solver *s, FN(orc_state) *st, const R *y) {
  FN(orc_hessian_diag)(s->nn, s->nt, s->tets, y, s->dminv, s->an, s->mu, s->lam, s->material, NULL, st->tethd);
#pragma omp parallel for schedule(static)
  for (int n = 0; n < s->nn; ++n) {
    R acc[3] = {0, 0, 0};
    const int e0 = s->inc_off[n], e1 = s->inc_off[n + 1];
    for (int q = 0; q < e1 - e0; ++q) {
      int slot = s->inc[s->p.gather_order ? e1 - 1 - q : e0 + q];
      for (int d = 0; d < 3; ++d) acc[d] += st->tethd[slot + d];
    }
    for (int d = 0; d < 3; ++d) st->hd[3 * n + d] = acc[d];
  }
}

/* Gradient of the incremental potential at y over free DOFs, given nodef = elastic force at y.
 * REF_GPU (CUDAInterface.cu:599-602): g = (y - x~) m/h^2 - f - f_ext - m*grav.
 * REF_CPU (DescentOptimize.cpp:43-61 with extForce of :101-102 and the force of
 * IsotropicNeohookeanMaterial.cpp:98 that already holds gravity+external): gravity and
 * external force enter twice.  Fixed nodes: g = 0.  Returns |g|^2 in FP64. */
static double FN(gradient)(orc_solver *s, FN(orc_state) *st, const R *y, const R *ext) {
  R h = (R)s->p.dt;
  R scale = s->p.semantics == ORC_REF_CPU ? (R)2 : (R)1;
  double g2 = 0;
#pragma omp parallel for schedule(static) reduction(+ : g2)
  for (int n = 0; n < s->nn; ++n) {
    if (s->mask[n]) {
      st->g[3 * n] = st->g[3 * n + 1] = st->g[3 * n + 2] = 0;
      continue;
    }
    R mh2 = (R)s->mass[n] / (h * h);
    for (int d = 0; d < 3; ++d) {
      R fe = (R)s->gforce[3 * n + d] + (ext ? ext[3 * n + d] : (R)0);
      R gd = (y[3 * n + d] - st->xt[3 * n + d]) * mh2 - st->nodef[3 * n + d] - scale * fe;
      st->g[3 * n + d] = gd;
      g2 += (double)gd * (double)gd;
    }
  }
  return g2;
}

/* Incremental potential at y (CUDAInterface.cu:587-597 + :542-550; DescentOptimize.cpp:63-89):
 * sum_t psi_t V_t + sum_free [ 1/2 m |y - x~|^2 / h^2 - m grav.y - f_ext.y ], summed in FP64. */
static double FN(energy)(orc_solver *s, FN(orc_state) *st, const R *y, const R *ext) {
  double E = 0;
  for (int t = 0; t < s->nt; ++t) E += (double)st->tete[t];
  R h = (R)s->p.dt;
  for (int n = 0; n < s->nn; ++n) {
    if (s->mask[n]) continue;
    R mh2 = (R)s->mass[n] / (h * h);
    R e = 0;
    for (int d = 0; d < 3; ++d) {
      R dd = y[3 * n + d] - st->xt[3 * n + d];
      e += (R)0.5 * dd * dd * mh2;
      e -= (ext ? ext[3 * n + d] : (R)0) * y[3 * n + d];
      e -= (R)s->gforce[3 * n + d] * y[3 * n + d];
    }
    E += (double)e;
  }
  return E;
}

static int FN(step)(orc_solver *s, const double *extd, const double *x0o, orc_stats *out, orc_trace *trace) {
  FN(orc_state) *st = (FN(orc_state) *)s->state;
  const orc_params *p = &s->p;
  int nn = s->nn;
  size_t n3 = 3 * (size_t)nn;
  R h = (R)p->dt;
  R *ext = NULL;
  if (extd) {
    ext = (R *)malloc(sizeof(R) * n3);
    for (size_t i = 0; i < n3; ++i) ext[i] = (R)extd[i];
  }
  R *xstart = (R *)malloc(sizeof(R) * n3);
  memcpy(xstart, st->x, sizeof(R) * n3);
  /* Inertial target and initial guess (CUDAInterface.cu:482-493; DescentOptimize.cpp:35-41). */
  for (size_t i = 0; i < n3; ++i) {
    st->xt[i] = st->x[i] + st->v[i] * h;
    if (p->semantics == ORC_REF_CPU)
      st->cur[i] = st->xt[i] + (R)(p->init_momentum * p->dt) * (st->v[i] - st->vl[i]);
    else
      st->cur[i] = (R)((double)st->xt[i] + (double)p->init_momentum * (double)(st->v[i] - st->vl[i]) * (double)h);
  }
  double g0 = -1;
  int used_learned = 0;
  if (x0o) {
    /* Learned guess: the plain guess's gradient is the denominator and the fallback test. */
    FN(forces_at)(s, st, st->cur);
    double g2p = FN(gradient)(s, st, st->cur, ext);
    R *cand = (R *)malloc(sizeof(R) * n3);
    for (size_t i = 0; i < n3; ++i) cand[i] = s->mask[i / 3] ? st->cur[i] : (R)x0o[i];
    FN(forces_at)(s, st, cand);
    double g2c = FN(gradient)(s, st, cand, ext);
    g0 = sqrt(g2p);
    if (g2c <= g2p) {
      memcpy(st->cur, cand, sizeof(R) * n3);
      used_learned = 1;
    }
    free(cand);
  }
  float alpha = p->alpha0;
  int k = 0, budget = p->iterations, passes = 0, rollbacks = 0, converged = 0;
  const int robust = p->relgrad_tol > 0;
  int kc = 0;
  float rho = s->rho, omega = s->omega;
  double E = 0, Eold = s->e_old, gl = 0;
  while (1) {
    if (alpha < p->alpha_min) break;
    if (!(k < budget)) break;
    int hdr = (k % p->hd_refresh == 0);
    if (hdr) FN(hd_at)(s, st, st->cur);
    FN(forces_at)(s, st, st->cur);
    double g2 = FN(gradient)(s, st, st->cur, ext);
    gl = sqrt(g2);
    if (g0 < 0) g0 = gl;
    int rb = 0, chk = 0;
    if (p->relgrad_tol > 0 && gl <= p->relgrad_tol * g0) {
      converged = 1;
      if (trace && passes < p->trace_cap) {
        orc_trace tr = {k, 0, hdr, 0, alpha, omega, E, gl};
        trace[passes] = tr;
      }
      passes++;
      break;
    }
    /* Jacobi step x_next = x - alpha g / (Hd + m/h^2) (CUDAInterface.cu:604-608). */
    for (int n = 0; n < nn; ++n) {
      R mh2 = (R)s->mass[n] / (h * h);
      for (int d = 0; d < 3; ++d) {
        int i = 3 * n + d;
        st->xnext[i] = s->mask[n] ? st->cur[i] : st->cur[i] - (R)alpha * st->g[i] / (st->hd[i] + mh2);
      }
    }
    if (k % p->energy_check == 0) {
      chk = 1;
      E = FN(energy)(s, st, st->cur, ext);
      /* Rel-grad mode (no reference counterpart) also rejects a non-finite energy. */
      int worse = robust ? !(E <= Eold) : (E > Eold);
      if (k != 0 && worse) {
        /* Rollback (CUDAInterface.cu:946-958): alpha *= 0.7, iterations -= k-8, redo k from x_old.
         * Rel-grad mode keeps the budget and restarts the Chebyshev profile at k. */
        alpha *= p->alpha_decay;
        if (!robust) budget -= k - p->energy_check;
        else kc = k;
        memcpy(st->cur, st->xold, sizeof(R) * n3);
        rollbacks++;
        rb = 1;
      } else {
        Eold = E;
        memcpy(st->xold, st->cur, sizeof(R) * n3);
      }
    }
    if (!rb) {
      /* Chebyshev schedule (CUDAInterface.cu:967-992; DescentOptimize.cpp:156-181) on the
       * iteration count since the last restart (kc = 0 in the reference). */
      int kk = k - kc;
      if (kk == 0) { rho = 0; omega = 1; }
      omega = 4.0f / (4.0f - rho * rho * omega);
      for (int i = 0; i < p->profile_n; ++i) {
        if (kk == p->profile_k[i] - 1) { rho = 0; omega = 1; }
        if (kk == p->profile_k[i]) { rho = p->profile_rho[i]; omega = 2.0f / (2.0f - rho * rho); break; }
      }
      if (omega != 1.0f)
        for (int n = 0; n < nn; ++n) {
          if (s->mask[n]) continue;
          for (int d = 0; d < 3; ++d) {
            int i = 3 * n + d;
            st->xnext[i] = (R)omega * (st->xnext[i] - st->xlast[i]) + st->xlast[i];
          }
        }
      memcpy(st->xlast, st->cur, sizeof(R) * n3);
      memcpy(st->cur, st->xnext, sizeof(R) * n3);
    }
    if (trace && passes < p->trace_cap) {
      orc_trace tr = {k, rb, hdr, chk, alpha, omega, E, gl};
      trace[passes] = tr;
    }
    passes++;
    if (!rb) k++;
  }
  s->rho = rho;
  s->omega = omega;
  s->e_old = Eold;
  /* Velocity update (CUDAInterface.cu:625-639 / DescentOptimize.cpp:183-185). */
  for (int n = 0; n < nn; ++n) {
    for (int d = 0; d < 3; ++d) {
      int i = 3 * n + d;
      if (p->semantics == ORC_REF_CPU) {
        st->vl[i] = st->v[i];
        st->v[i] = (R)p->damping * (st->cur[i] - xstart[i]) / h;
      } else if (!s->mask[n]) {
        st->vl[i] = st->v[i];
        st->v[i] += (st->cur[i] - st->xt[i]) * ((R)1 / h);
        st->v[i] = (R)((double)st->v[i] * (double)p->damping);
      }
    }
  }
  memcpy(st->x, st->cur, sizeof(R) * n3);
  out->passes = passes;
  out->iterations = k;
  out->rollbacks = rollbacks;
  out->converged = converged;
  out->used_learned = used_learned;
  out->alpha = alpha;
  out->energy = Eold;
  out->g0_norm = g0;
  out->g_norm = gl;
  free(xstart);
  free(ext);
  return 0;
}

static double FN(grad_norm_at)(orc_solver *s, const double *xd) {
  FN(orc_state) *st = (FN(orc_state) *)s->state;
  size_t n3 = 3 * (size_t)s->nn;
  R *y = (R *)malloc(sizeof(R) * n3);
  for (size_t i = 0; i < n3; ++i) y[i] = (R)xd[i];
  FN(forces_at)(s, st, y);
  double g2 = FN(gradient)(s, st, y, NULL);
  free(y);
  return sqrt(g2);
}

#undef FN
#undef ORC_CAT
#undef ORC_CAT_
