// pbs_device.h -- per-tet device math of the descent solver, templated on the real type R
// (float: the reference's FP32 arithmetic; double: FP64 arithmetic on FP32 mesh data).
//
// Conventions (same as the reference's GPU path, CUDAInterface.cu:495-568): 3x3 matrices are
// row-major; Ds = [x1-x0, x2-x0, x3-x0] (columns); F = Ds Dm^-1; the elastic force on local
// nodes 1..3 is the matching column of G = P AN and node 0 gets minus their sum.
// AN = -vol Dm^-T: for a positively oriented rest tet this equals the reference's
// -(v1 x v2 + v2 x v3 + v3 x v1)/6 area vectors (TetMesh.cpp:279-294), so it is recomputed in
// registers instead of being stored.
#pragma once
#include <hip/hip_runtime.h>

namespace pbs {

// FP64 natural logarithm in about 40 instructions (the library's takes about 75, 15 % of the FP64
// force evaluation).  x = 2^e m with m in [sqrt(1/2), sqrt(2)); f = m - 1 (exact); s = f / (2 + f)
// from the hardware reciprocal refined by two Newton steps and a residual correction;
// log(1 + f) = f - (f^2/2 - s (f^2/2 + R)), R = sum_{k=1..9} 2 z^k / (2k + 1), z = s^2: the series
// of log((1 + s) / (1 - s)) = 2 atanh(s) in the evaluation form of Sun's fdlibm (e_log.c), with the
// series' own coefficients (|z| <= 0.0295, so the dropped terms stay below 2.3e-17 relative);
// then + e ln 2 in two parts (the high part has 32 significant bits, so e * hi is exact).  Within
// 1 ulp of the correctly rounded logarithm (80 M samples on [0.5, 2], around 1 and over the whole
// range); 0 gives -inf, +inf gives +inf, negative numbers and NaN give NaN, as log() does.
// The series coefficients are materialised in scalar registers at each use (an empty asm statement
// the compiler may not hoist): left to itself, the compiler keeps all nine in 18 VGPRs across the
// tet loop (FP64 tile pass: 94 instead of 85 VGPRs).  The arithmetic is the same either way.
__device__ __forceinline__ double log_f64(double x) {
  double m = __builtin_amdgcn_frexp_mant(x);  // [0.5, 1)
  int e = __builtin_amdgcn_frexp_exp(x);
  if (m < 0.70710678118654752440) {
    m = 2.0 * m;
    e -= 1;
  }
  const double f = m - 1.0, d = 2.0 + f;
  double r = __builtin_amdgcn_rcp(d);
  double t = fma(-d, r, 1.0);
  r = fma(r, t, r);
  t = fma(-d, r, 1.0);
  r = fma(r, t, r);
  double s = f * r;
  t = fma(-s, d, f);
  s = fma(t, r, s);
  const double z = s * s;
  auto c = [](double v) {
    asm volatile("" : "+s"(v));
    return v;
  };
  double R = c(2.0 / 19.0);
  R = fma(R, z, c(2.0 / 17.0));
  R = fma(R, z, c(2.0 / 15.0));
  R = fma(R, z, c(2.0 / 13.0));
  R = fma(R, z, c(2.0 / 11.0));
  R = fma(R, z, c(2.0 / 9.0));
  R = fma(R, z, c(2.0 / 7.0));
  R = fma(R, z, c(2.0 / 5.0));
  R = fma(R, z, c(2.0 / 3.0));
  R *= z;
  const double hf = 0.5 * f * f;
  const double l1 = f - (hf - s * (hf + R));
  const double de = (double)e;
  double res = fma(de, 0.6931471803691238, fma(de, 1.9082149292705877e-10, l1));
  if (!(x > 0.0)) res = x == 0.0 ? -__builtin_inf() : __builtin_nan("");
  if (x == __builtin_inf()) res = x;
  return res;
}

template <typename R> __device__ __forceinline__ R dlog(R v);
template <> __device__ __forceinline__ float dlog<float>(float v) { return logf(v); }
template <> __device__ __forceinline__ double dlog<double>(double v) { return log_f64(v); }
template <typename R> __device__ __forceinline__ R drsqrt(R v);
template <> __device__ __forceinline__ float drsqrt<float>(float v) { return rsqrtf(v); }
template <> __device__ __forceinline__ double drsqrt<double>(double v) { return rsqrt(v); }
template <typename R> __device__ __forceinline__ R dsqrt(R v);
template <> __device__ __forceinline__ float dsqrt<float>(float v) { return sqrtf(v); }
template <> __device__ __forceinline__ double dsqrt<double>(double v) { return sqrt(v); }

// Adjugate of A (row-major) and det(A) expanded down column 0 (cuMath::Matrix_Inverse_3).
template <typename R> __device__ __forceinline__ R adj3(const R *A, R *C) {
  C[0] = A[4] * A[8] - A[7] * A[5];
  C[1] = A[7] * A[2] - A[1] * A[8];
  C[2] = A[1] * A[5] - A[4] * A[2];
  C[3] = A[5] * A[6] - A[3] * A[8];
  C[4] = A[0] * A[8] - A[2] * A[6];
  C[5] = A[2] * A[3] - A[0] * A[5];
  C[6] = A[3] * A[7] - A[4] * A[6];
  C[7] = A[1] * A[6] - A[0] * A[7];
  C[8] = A[0] * A[4] - A[1] * A[3];
  return A[0] * C[0] + A[3] * C[1] + A[6] * C[2];
}

template <typename R> __device__ __forceinline__ void mul3(const R *A, const R *B, R *C) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

// ---------------------------------------------------------------------------------------------
// 3x3 SVD after McAdams, Selle, Tamstorf, Teran, Sifakis, "Computing the Singular Value
// Decomposition of 3x3 matrices with minimal branching and elementary floating point
// operations", UW-Madison CS TR1690 (2011): Jacobi eigenanalysis of F^T F (here with exact
// Jacobi rotations and a fixed sweep count), columns of B = F V sorted by norm with sign-
// preserving swaps (det V = +1), then a Givens QR of B giving U (det U = +1) and the singular
// values on the diagonal of R; the last one carries the sign of det F.
// ---------------------------------------------------------------------------------------------
template <typename R> struct SvdSweeps;
template <> struct SvdSweeps<float> { static constexpr int n = 5; };
template <> struct SvdSweeps<double> { static constexpr int n = 8; };

template <typename R> __device__ __forceinline__ void jacobi_rot(R *A, R *V, int p, int q, int r) {
  // A symmetric stored full row-major; rotate in the (p, q) plane, r is the third index.
  R apq = A[3 * p + q];
  R app = A[3 * p + p], aqq = A[3 * q + q];
  R tiny = (sizeof(R) == 4 ? (R)1e-20 : (R)1e-150);
  if (fabs(apq) <= tiny * (fabs(app) + fabs(aqq)) + tiny) return;
  R th = (aqq - app) / ((R)2 * apq);
  R t = (th >= (R)0 ? (R)1 : (R)-1) / (fabs(th) + dsqrt(th * th + (R)1));
  R c = drsqrt(t * t + (R)1), s = t * c;
  A[3 * p + p] = app - t * apq;
  A[3 * q + q] = aqq + t * apq;
  A[3 * p + q] = A[3 * q + p] = (R)0;
  R arp = A[3 * r + p], arq = A[3 * r + q];
  R nrp = c * arp - s * arq, nrq = s * arp + c * arq;
  A[3 * r + p] = A[3 * p + r] = nrp;
  A[3 * r + q] = A[3 * q + r] = nrq;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    R vp = V[3 * i + p], vq = V[3 * i + q];
    V[3 * i + p] = c * vp - s * vq;
    V[3 * i + q] = s * vp + c * vq;
  }
}

template <typename R> __device__ __forceinline__ void givens_qr_step(R *B, R *U, int p, int q) {
  // zero B[q][p] using rows p and q (p < q): rows' = G rows, U = U G^T
  R a = B[3 * p + p], b = B[3 * q + p];
  R rr = a * a + b * b;
  R c = (R)1, s = (R)0;
  if (rr > (R)0) {
    R inv = drsqrt(rr);
    c = a * inv;
    s = b * inv;
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    R bp = B[3 * p + j], bq = B[3 * q + j];
    B[3 * p + j] = c * bp + s * bq;
    B[3 * q + j] = -s * bp + c * bq;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    R up = U[3 * i + p], uq = U[3 * i + q];
    U[3 * i + p] = c * up + s * uq;
    U[3 * i + q] = -s * up + c * uq;
  }
}

template <typename R> __device__ void svd3(const R *F, R *U, R *S, R *V) {
  R A[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) A[3 * i + j] = F[i] * F[j] + F[3 + i] * F[3 + j] + F[6 + i] * F[6 + j];
#pragma unroll
  for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? (R)1 : (R)0;
  for (int sw = 0; sw < SvdSweeps<R>::n; ++sw) {
    jacobi_rot(A, V, 0, 1, 2);
    jacobi_rot(A, V, 0, 2, 1);
    jacobi_rot(A, V, 1, 2, 0);
  }
  R B[9];
  mul3(F, V, B);
  R n0 = B[0] * B[0] + B[3] * B[3] + B[6] * B[6];
  R n1 = B[1] * B[1] + B[4] * B[4] + B[7] * B[7];
  R n2 = B[2] * B[2] + B[5] * B[5] + B[8] * B[8];
  // sort columns by descending norm; each swap negates one column so det V stays +1
  auto swapcols = [&](int a, int b) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      R tb = B[3 * i + a];
      B[3 * i + a] = -B[3 * i + b];
      B[3 * i + b] = tb;
      R tv = V[3 * i + a];
      V[3 * i + a] = -V[3 * i + b];
      V[3 * i + b] = tv;
    }
  };
  if (n0 < n1) { swapcols(0, 1); R t = n0; n0 = n1; n1 = t; }
  if (n0 < n2) { swapcols(0, 2); R t = n0; n0 = n2; n2 = t; }
  if (n1 < n2) { swapcols(1, 2); R t = n1; n1 = n2; n2 = t; }
  (void)n0; (void)n1; (void)n2;
#pragma unroll
  for (int i = 0; i < 9; ++i) U[i] = (i % 4 == 0) ? (R)1 : (R)0;
  givens_qr_step(B, U, 0, 1);
  givens_qr_step(B, U, 0, 2);
  givens_qr_step(B, U, 1, 2);
  S[0] = B[0];
  S[1] = B[4];
  S[2] = B[8];
}

// ---------------------------------------------------------------------------------------------
// Per-tet kernels' body: forces, energy and the Hessian (stiffness) diagonal.
// ---------------------------------------------------------------------------------------------
// Node-n gradient-of-shape vectors c_n (rows of Dm^-1, node 0 = minus their sum).
template <typename R> __device__ __forceinline__ void shape_vec(const R *Dmi, int n, R *c) {
  if (n == 0) {
#pragma unroll
    for (int j = 0; j < 3; ++j) c[j] = -(Dmi[j] + Dmi[3 + j] + Dmi[6 + j]);
  } else {
#pragma unroll
    for (int j = 0; j < 3; ++j) c[j] = Dmi[3 * (n - 1) + j];
  }
}

// Neo-Hookean force / energy / stiffness diagonal.
//   P = mu (F - F^-T) + lambda ln J F^-T                    (CUDAInterface.cu:221-232)
//   psi = mu/2 (|F|^2 - 3) - mu ln J + lambda/2 (ln J)^2  (CUDAInterface.cu:542-549)
//   K(n,d) = vol (mu |c_n|^2 + (mu + lambda - lambda ln J) (c_n . w_d)^2), w_d = column d of F^-1:
//   the diagonal of c_n^T (dP_d./dF_d.) c_n for the closed-form derivative
//   dP_ij/dF_kl = mu d_ik d_jl + (mu - lambda ln J) Finv_li Finv_jk + lambda Finv_ji Finv_lk,
//   which is what the reference's SVD/Xu-2015 route (CUDAInterface.cu:101-219) evaluates.
template <typename R, bool HD>
__device__ __forceinline__ void neo_tet(const R *F, const R *Dmi, R vol, R mu, R lam, R *f12, R &energy, R *h12) {
  R C[9];
  R J = adj3(F, C);
  R iJ = (R)1 / J;
  R Fi[9];  // F^-1 = C / J
#pragma unroll
  for (int i = 0; i < 9; ++i) Fi[i] = C[i] * iJ;
  R lj = dlog(J);
  R P[9];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      R fit = Fi[3 * c + a];  // (F^-T)[a][c]
      P[3 * a + c] = (F[3 * a + c] - fit) * mu + lam * lj * fit;
    }
  // G = P AN = -vol P Dm^-T: G[a][b] = -vol sum_c P[a][c] Dmi[b][c]
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    R g0 = -vol * (P[3 * a] * Dmi[0] + P[3 * a + 1] * Dmi[1] + P[3 * a + 2] * Dmi[2]);
    R g1 = -vol * (P[3 * a] * Dmi[3] + P[3 * a + 1] * Dmi[4] + P[3 * a + 2] * Dmi[5]);
    R g2 = -vol * (P[3 * a] * Dmi[6] + P[3 * a + 1] * Dmi[7] + P[3 * a + 2] * Dmi[8]);
    f12[3 + a] = g0;
    f12[6 + a] = g1;
    f12[9 + a] = g2;
    f12[a] = -(g0 + g1 + g2);
  }
  R I1 = (R)0;
#pragma unroll
  for (int i = 0; i < 9; ++i) I1 += F[i] * F[i];
  energy = ((R)0.5 * mu * (I1 - (R)3) - mu * lj + (R)0.5 * lam * lj * lj) * vol;
  if (HD) {
    R k2 = mu + lam - lam * lj;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      R c[3];
      shape_vec(Dmi, n, c);
      R cc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        R s = c[0] * Fi[d] + c[1] * Fi[3 + d] + c[2] * Fi[6 + d];
        h12[3 * n + d] = vol * (mu * cc + k2 * s * s);
      }
    }
  }
}

// Neo-Hookean forces from the displacement gradient H = F - I (F = I + H), free of the
// cancellation F ~ I causes in FP32: with t = tr H, s2 = ((tr H)^2 - tr(H^2)) / 2 and
// delta = J - 1 = t + s2 + det H, Cayley-Hamilton gives adj(I + H) = (1 + t + s2) I - (1 + t) H + H^2,
// so E = F^-1 - I = (-det H I - (1 + t) H + H^2) / J and
//   P = mu (H - E^T) + lambda ln(1 + delta) (I + E^T),
// every term a product of small quantities.  Forces as neo_tet: G = -vol P Dm^-T.
template <typename R>
__device__ __forceinline__ void neo_tet_disp(const R *H, const R *Dmi, R vol, R mu, R lam, R *f12) {
  const R t = H[0] + H[4] + H[8];
  R H2[9];
  mul3(H, H, H2);
  const R s2 = (R)0.5 * (t * t - (H2[0] + H2[4] + H2[8]));
  R Cf[9];
  const R dH = adj3(H, Cf);
  const R delta = t + s2 + dH;
  const R iJ = (R)1 / ((R)1 + delta);
  const R lj = (sizeof(R) == 4) ? (R)log1pf((float)delta) : (R)log1p((double)delta);
  const R c1 = (R)1 + t;
  R P[9];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      // E^T[a][c] = E[c][a]
      const R et = ((c == a ? -dH : (R)0) - c1 * H[3 * c + a] + H2[3 * c + a]) * iJ;
      P[3 * a + c] = mu * (H[3 * a + c] - et) + lam * lj * ((c == a ? (R)1 : (R)0) + et);
    }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    R g0 = -vol * (P[3 * a] * Dmi[0] + P[3 * a + 1] * Dmi[1] + P[3 * a + 2] * Dmi[2]);
    R g1 = -vol * (P[3 * a] * Dmi[3] + P[3 * a + 1] * Dmi[4] + P[3 * a + 2] * Dmi[5]);
    R g2 = -vol * (P[3 * a] * Dmi[6] + P[3 * a + 1] * Dmi[7] + P[3 * a + 2] * Dmi[8]);
    f12[3 + a] = g0;
    f12[6 + a] = g1;
    f12[9 + a] = g2;
    f12[a] = -(g0 + g1 + g2);
  }
}

// Energy alone (psi vol, as neo_tet / stretch_tet compute it), for the mixed mode's check passes.
template <typename R, int MAT> __device__ __forceinline__ R tet_energy(const R *F, R vol, R mu, R lam) {
  R C[9];
  const R J = adj3(F, C);
  if (MAT == 0) {
    R I1 = (R)0;
#pragma unroll
    for (int i = 0; i < 9; ++i) I1 += F[i] * F[i];
    const R lj = dlog(J);
    return ((R)0.5 * mu * (I1 - (R)3) - mu * lj + (R)0.5 * lam * lj * lj) * vol;
  } else {
    R U[9], S[3], V[9];
    svd3(F, U, S, V);
    const R e = (S[0] - (R)1) * (S[0] - (R)1) + (S[1] - (R)1) * (S[1] - (R)1) + (S[2] - (R)1) * (S[2] - (R)1);
    return (mu * e + (R)0.5 * lam * (J - (R)1) * (J - (R)1)) * vol;
  }
}

// Rotation of the polar decomposition F = R P for det F > 0 by Higham's scaled Newton iteration,
// X <- (zeta X + X^-T / zeta) / 2 with zeta = sqrt(|X^-1|_F / |X|_F) (N. J. Higham, "Computing the
// polar decomposition with applications", SIAM J. Sci. Stat. Comput. 7, 1986).  With
// kappa = |F|_F |F^-1|_F <= 20, 4 iterations (FP32) / 5 (FP64) reach the working precision (largest
// error against U V^T over 2e4 random F: 3e-7 / 1.4e-14).  Returns false (Rot untouched) when
// det F <= 0 or kappa is larger.
template <typename R> __device__ __forceinline__ bool polar_rotation(const R *F, R *Rot) {
  constexpr int kIters = sizeof(R) == 4 ? 4 : 5;
  R X[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) X[i] = F[i];
  bool ok = true;
#pragma unroll
  for (int it = 0; it < kIters; ++it) {
    R C[9];
    const R det = adj3(X, C);  // X^-1 = C / det, (X^-T)[a][c] = C[3c + a] / det
    R nx = (R)0, nc = (R)0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      nx += X[i] * X[i];
      nc += C[i] * C[i];
    }
    if (it == 0) ok = det > (R)0 && nc * nx <= (R)400 * det * det;
    const R zeta = sqrt(sqrt(nc / nx) / fabs(det));
    const R a = (R)0.5 * zeta, b = (R)0.5 / (zeta * det);
    R Y[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) Y[3 * r + c] = a * X[3 * r + c] + b * C[3 * c + r];
#pragma unroll
    for (int i = 0; i < 9; ++i) X[i] = Y[i];
  }
  if (ok)
#pragma unroll
    for (int i = 0; i < 9; ++i) Rot[i] = X[i];
  return ok;
}

// Stretch material psi = sum mu (s_i - 1)^2 + lambda/2 (J - 1)^2 on signed singular values.
//   P = 2 mu (F - U V^T) + lambda (J - 1) J F^-T.
//   K(n,d) = vol [ 2 mu (|c_n|^2 - sum_{k<l} (a_kl . c_n)^2 / (s_k + s_l)) + lambda J^2 (c_n . w_d)^2 ]
//   with a_kl[c] = U_dk V_cl - U_dl V_ck (the rotation part of Xu et al. 2015's 2x2 solves,
//   written in closed form).
template <typename R, bool HD>
__device__ __forceinline__ void stretch_tet(const R *F, const R *Dmi, R vol, R mu, R lam, R *f12, R &energy,
                                            R *h12) {
  // Forces and energy need only R = U V^T: for det F > 0 it is the polar rotation, and
  // sum (s_i - 1)^2 = |F - R|_F^2.  The Hessian diagonal and inverted or badly conditioned tets
  // take the SVD.
  R U[9], S[3], V[9], Rot[9];
  const bool polar = !HD && polar_rotation(F, Rot);
  if (!polar) {
    svd3(F, U, S, V);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        Rot[3 * i + j] = U[3 * i] * V[3 * j] + U[3 * i + 1] * V[3 * j + 1] + U[3 * i + 2] * V[3 * j + 2];
  }
  R C[9];
  R J = adj3(F, C);
  R iJ = (R)1 / J;
  R Fi[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) Fi[i] = C[i] * iJ;
  R cj = lam * (J - (R)1) * J;
  R P[9];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) P[3 * a + c] = (R)2 * mu * (F[3 * a + c] - Rot[3 * a + c]) + cj * Fi[3 * c + a];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    R g0 = -vol * (P[3 * a] * Dmi[0] + P[3 * a + 1] * Dmi[1] + P[3 * a + 2] * Dmi[2]);
    R g1 = -vol * (P[3 * a] * Dmi[3] + P[3 * a + 1] * Dmi[4] + P[3 * a + 2] * Dmi[5]);
    R g2 = -vol * (P[3 * a] * Dmi[6] + P[3 * a + 1] * Dmi[7] + P[3 * a + 2] * Dmi[8]);
    f12[3 + a] = g0;
    f12[6 + a] = g1;
    f12[9 + a] = g2;
    f12[a] = -(g0 + g1 + g2);
  }
  R e = (R)0;
  if (polar) {
#pragma unroll
    for (int i = 0; i < 9; ++i) e += (F[i] - Rot[i]) * (F[i] - Rot[i]);
  } else {
    e = (S[0] - (R)1) * (S[0] - (R)1) + (S[1] - (R)1) * (S[1] - (R)1) + (S[2] - (R)1) * (S[2] - (R)1);
  }
  energy = (mu * e + (R)0.5 * lam * (J - (R)1) * (J - (R)1)) * vol;
  if (HD) {
    const int kk[3] = {0, 0, 1}, ll[3] = {1, 2, 2};
    R inv[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      R den = S[kk[q]] + S[ll[q]];
      R lim = (sizeof(R) == 4 ? (R)1e-6 : (R)1e-12);
      inv[q] = (R)1 / (den > lim ? den : lim);
    }
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      R c[3];
      shape_vec(Dmi, n, c);
      R cc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        R rot = (R)0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          int k = kk[q], l = ll[q];
          R ac = (U[3 * d + k] * V[l] - U[3 * d + l] * V[k]) * c[0] +
                 (U[3 * d + k] * V[3 + l] - U[3 * d + l] * V[3 + k]) * c[1] +
                 (U[3 * d + k] * V[6 + l] - U[3 * d + l] * V[6 + k]) * c[2];
          rot += ac * ac * inv[q];
        }
        R s = c[0] * Fi[d] + c[1] * Fi[3 + d] + c[2] * Fi[6 + d];
        h12[3 * n + d] = vol * ((R)2 * mu * (cc - rot) + lam * J * J * s * s);
      }
    }
  }
}

}  // namespace pbs
