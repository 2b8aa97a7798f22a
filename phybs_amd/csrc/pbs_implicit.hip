// pbs_implicit.hip -- kernels of the backward-Euler step (computeInnerforces,
// computeGlobalStiffnessMatrix and doBackEuler, CUDAInterface.cu:784-884, as driven by
// fvm.cpp:179-183).
//
// The reference scatters 144 float atomics per tet into a CSR matrix (computeGlobalK,
// CUDAInterface.cu:234-271) and solves with a sparse Cholesky factorisation.  Here one thread
// owns one block row (a node): it walks the node's incident tets in a fixed order, evaluates the
// 3x12 stiffness strip of its corner in closed form and accumulates it into its own SELL-64
// slots, so the assembly is deterministic and atomic-free; the solve is a Jacobi-preconditioned
// conjugate gradient in FP64 whose scalars stay on the device.
#include <hip/hip_ext.h>

#include "pbs_device.h"
#include "pbs_implicit.h"

namespace pbs {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ size_t val_idx(long group, int q, int lane) {
  return ((size_t)group * 9 + q) * 64 + lane;
}

// XCD-contiguous block order: workgroup i runs on XCD i % 8; giving each XCD a contiguous range
// of row blocks keeps a row's neighbours (mostly in the same or the adjacent blocks, in Morton
// order) in that XCD's L2.
__device__ __forceinline__ int xcd_block(int bid, int nblk) {
  const int per = (nblk + 7) >> 3;
  const int xcd = bid & 7, k = bid >> 3;
  const int full = nblk - 8 * (per - 1);
  const int start = xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1);
  return start + k;
}

// Sum of v over the 256 threads of the block in a fixed order (wave tree, then the 4 waves).
__device__ __forceinline__ double block_sum256(double v, double *lds) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) lds[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) t = (lds[0] + lds[1]) + (lds[2] + lds[3]);
  __syncthreads();
  return t;
}

// Sum of component c of the per-block partials, by one block of 1024 threads in a fixed order.
template <int K>
__device__ __forceinline__ void partial_sums(const double *part, int nb, const int (&c)[K], double *lds, double (&out)[K]) {
  // per thread: part[4 i + c] for i = tid, tid + 1024, ... in that order, the loads of 8 terms
  // issued before their additions (indices past the end clamped to the first, not added)
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = 0.0;
  const int first = threadIdx.x;
  for (int base = first; base < nb; base += 8 * 1024) {
    double t[8][K];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = base + u * 1024 < nb ? base + u * 1024 : first;
#pragma unroll
      for (int k = 0; k < K; ++k) t[u][k] = part[4 * (size_t)i + c[k]];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (base + u * 1024 < nb)
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] += t[u][k];
  }
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_down(v[k], o, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) lds[16 * k + w] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double t = 0.0;
    if (threadIdx.x == 0)
      for (int i = 0; i < 16; ++i) t += lds[16 * k + i];
    out[k] = t;
  }
  __syncthreads();
}

// y += (row n of A) v.  BLK = 3: 3x3 blocks (9 values per slot); BLK = 1: one value per slot,
// the block being that value times the 3x3 identity.  get(col, v0, v1, v2) supplies v at a column.
template <typename R, int BLK, typename Get>
__device__ __forceinline__ void row_apply(const IeDev &d, int n, Get get, double *y) {
  const R *__restrict__ A = (const R *)d.A;
  // all lanes run to the slice width (padding slots hold zeros and the row's own column), so the
  // loop is uniform across the wavefront and can be unrolled to overlap its loads
  const int lane = n & 63, g0 = d.slice_off[n >> 6], width = d.slice_off[(n >> 6) + 1] - g0;
#pragma unroll 2
  for (int j = 0; j < width; ++j) {
    const int col = d.bcol[(size_t)(g0 + j) * 64 + lane];
    double x0, x1, x2;
    get(col, x0, x1, x2);
    if (BLK == 3) {
#pragma unroll
      for (int q = 0; q < 3; ++q)
        y[q] += (double)A[val_idx(g0 + j, 3 * q, lane)] * x0 + (double)A[val_idx(g0 + j, 3 * q + 1, lane)] * x1 +
                (double)A[val_idx(g0 + j, 3 * q + 2, lane)] * x2;
    } else {
      const double a = (double)A[(size_t)(g0 + j) * 64 + lane];
      y[0] += a * x0;
      y[1] += a * x1;
      y[2] += a * x2;
    }
  }
}

// Slot (within the row) of column `col`: the row's columns are ascending.
__device__ __forceinline__ int find_col(const int *__restrict__ bcol, int g0, int lane, int len, int col) {
  int lo = 0, hi = len - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (bcol[(size_t)(g0 + mid) * 64 + lane] < col) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// ---------------------------------------------------------------------------------------------
// Assembly.  For node n at corner i of an incident tet, with shape vectors c_m (rows of Dm^-1,
// c_0 = -(c_1 + c_2 + c_3)), g_m = F^-T c_m and V the rest volume, the stiffness block
// K(i,j)[d][e] = -d f_i[d] / d x_j[e] is
//   neo-Hookean: V [ mu (c_i . c_j) delta_de + (mu - lambda ln J) g_j[d] g_i[e] + lambda g_i[d] g_j[e] ]
//   stretch:     V [ 2 mu ((c_i . c_j) delta_de - sum_{k<l} (a^d_kl . c_i)(a^e_kl . c_j) / (s_k + s_l))
//                    + lambda ((2J - 1) J g_i[d] g_j[e] - (J - 1) J g_i[e] g_j[d]) ]
// with a^d_kl[b] = U_dk V_bl - U_dl V_bk (F = U diag(s) V^T): the closed form of the
// derivative the reference evaluates through the SVD and Xu et al. 2015's 2x2 solves
// (compute_dfdx, CUDAInterface.cu:101-219).  The elastic force of the corner comes from the same
// per-tet routine as the descent solver's.
// LACC: the row (`row`) is accumulated in LDS (one 64-row slice per workgroup, each lane its own
// column of slots, so no lane touches another's) and written to the matrix once; otherwise it is
// accumulated in place in the matrix.
// ---------------------------------------------------------------------------------------------
template <typename R, int MAT, bool LACC>
__global__ void __launch_bounds__(kBlock)
    k_ie_assemble(DevParams p, IeAsm a, IeDev d, const int4 *__restrict__ tets, const float *__restrict__ dmi,
                  const float *__restrict__ vol, const int *__restrict__ inc_off, const int *__restrict__ inc,
                  const R *__restrict__ xpos, const R *__restrict__ vel, const float *__restrict__ rest,
                  const float *__restrict__ mass, const float *__restrict__ gforce, const float *__restrict__ ext,
                  const int *__restrict__ mask, const float *__restrict__ inst_mu, const float *__restrict__ inst_lam) {
  extern __shared__ __align__(16) unsigned char ie_lds[];
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= d.n) return;
  R *__restrict__ A = (R *)d.A;
  const int N = p.nn, T = p.nt;
  const int lane = n & 63;
  const int g0 = d.slice_off[n >> 6];
  const int width = d.slice_off[(n >> 6) + 1] - g0;
  const int len = d.row_len[n];
  R *__restrict__ row = LACC ? (R *)ie_lds : A;
  auto ix = [&](int j, int q) -> size_t { return LACC ? ((size_t)j * 9 + q) * 64 + lane : val_idx(g0 + j, q, lane); };
  for (int j = 0; j < width; ++j)
#pragma unroll
    for (int q = 0; q < 9; ++q) row[ix(j, q)] = (R)0;
  const R mu = (R)inst_mu[0], lam = (R)inst_lam[0];
  R f[3] = {(R)0, (R)0, (R)0};
  for (int e = inc_off[n]; e < inc_off[n + 1]; ++e) {
    const int t = inc[e] >> 2, ci = inc[e] & 3;
    const int4 q4 = tets[t];
    const int qn[4] = {q4.x, q4.y, q4.z, q4.w};
    R Dmi[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Dmi[k] = (R)dmi[(size_t)k * T + t];
    const R V = (R)vol[t];
    R X[4][3];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int k = 0; k < 3; ++k) X[c][k] = xpos[(size_t)k * N + qn[c]];
    R Ds[9], F[9];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      Ds[3 * k + 0] = X[1][k] - X[0][k];
      Ds[3 * k + 1] = X[2][k] - X[0][k];
      Ds[3 * k + 2] = X[3][k] - X[0][k];
    }
    mul3(Ds, Dmi, F);
    R f12[12], en;
    if (MAT == 0) neo_tet<R, false>(F, Dmi, V, mu, lam, f12, en, nullptr);
    else stretch_tet<R, false>(F, Dmi, V, mu, lam, f12, en, nullptr);
#pragma unroll
    for (int k = 0; k < 3; ++k) f[k] += f12[3 * ci + k];
    R C[9];
    const R J = adj3(F, C);
    const R iJ = (R)1 / J;
    R c[4][3], g[4][3];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      shape_vec(Dmi, m, c[m]);
#pragma unroll
      for (int k = 0; k < 3; ++k) g[m][k] = (c[m][0] * C[k] + c[m][1] * C[3 + k] + c[m][2] * C[6 + k]) * iJ;
    }
    R ci3[3], gi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {  // corner ci's vectors, selected without dynamic register indexing
      ci3[k] = ci == 0 ? c[0][k] : ci == 1 ? c[1][k] : ci == 2 ? c[2][k] : c[3][k];
      gi[k] = ci == 0 ? g[0][k] : ci == 1 ? g[1][k] : ci == 2 ? g[2][k] : g[3][k];
    }
    R k1, k2;          // coefficients of g_j[d] g_i[e] and g_i[d] g_j[e]
    R ai[3][3] = {};   // stretch: (a^d_kl . c_i) for the pairs kl = 01, 02, 12
    R aw[3] = {};      // stretch: 1 / (s_k + s_l)
    R U[9] = {}, Vv[9] = {};
    if (MAT == 0) {
      const R lj = dlog(J);
      k1 = mu - lam * lj;
      k2 = lam;
    } else {
      R S[3];
      svd3(F, U, S, Vv);
      k1 = -lam * (J - (R)1) * J;
      k2 = lam * ((R)2 * J - (R)1) * J;
      const int kk[3] = {0, 0, 1}, ll[3] = {1, 2, 2};
      const R lim = (sizeof(R) == 4 ? (R)1e-6 : (R)1e-12);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const R den = S[kk[q]] + S[ll[q]];
        aw[q] = (R)1 / (den > lim ? den : lim);
#pragma unroll
        for (int dd = 0; dd < 3; ++dd) {
          R acc = (R)0;
#pragma unroll
          for (int b = 0; b < 3; ++b)
            acc += (U[3 * dd + kk[q]] * Vv[3 * b + ll[q]] - U[3 * dd + ll[q]] * Vv[3 * b + kk[q]]) * ci3[b];
          ai[q][dd] = acc;
        }
      }
    }
    const R mu2 = (MAT == 0) ? mu : (R)2 * mu;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int slot_j = find_col(d.bcol, g0, lane, len, qn[m]);
      const R cc = ci3[0] * c[m][0] + ci3[1] * c[m][1] + ci3[2] * c[m][2];
      R aj[3][3];
      if (MAT != 0) {
        const int kk[3] = {0, 0, 1}, ll[3] = {1, 2, 2};
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
          for (int ee = 0; ee < 3; ++ee) {
            R acc = (R)0;
#pragma unroll
            for (int b = 0; b < 3; ++b)
              acc += (U[3 * ee + kk[q]] * Vv[3 * b + ll[q]] - U[3 * ee + ll[q]] * Vv[3 * b + kk[q]]) * c[m][b];
            aj[q][ee] = acc;
          }
      }
#pragma unroll
      for (int dd = 0; dd < 3; ++dd)
#pragma unroll
        for (int ee = 0; ee < 3; ++ee) {
          R kv = (dd == ee ? mu2 * cc : (R)0) + k1 * g[m][dd] * gi[ee] + k2 * gi[dd] * g[m][ee];
          if (MAT != 0) {
            R rot = (R)0;
#pragma unroll
            for (int q = 0; q < 3; ++q) rot += ai[q][dd] * aj[q][ee] * aw[q];
            kv -= mu2 * rot;
          }
          row[ix(slot_j, 3 * dd + ee)] += V * kv;
        }
    }
  }
  auto flush = [&]() {  // LDS row -> matrix
    if (LACC)
      for (int j = 0; j < width; ++j)
#pragma unroll
        for (int q = 0; q < 9; ++q) A[val_idx(g0 + j, q, lane)] = row[ix(j, q)];
  };
  if (a.raw) {
    flush();
    return;
  }
  // system matrix (h + beta) h K + (1 + alpha h) M (doBackEuler, CUDAInterface.cu:845-851)
  const R s1 = (R)a.h + (R)a.beta, s2 = (R)a.h;
  for (int j = 0; j < len; ++j)
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      const size_t i = ix(j, q);
      row[i] = row[i] * s1 * s2;
    }
  const R ms = (R)mass[n] * ((R)1 + (R)a.alpha * (R)a.h);
  const int ds = d.diag_slot[n];
  const int dj = (int)((ds >> 6) - g0);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const size_t i = ix(dj, 4 * k);
    row[i] += ms;
    d.dinv[3 * (size_t)n + k] = 1.0 / (double)row[i];
  }
  flush();
  // right-hand side h f + m v (setRHSofLinearSystem :418-442) with the forces of
  // addFixContraintsAndGravity (:327-343): fixed nodes a 1e8 spring to the rest position
  const R h = (R)a.h;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const size_t i = (size_t)k * N + n;
    R fk;
    if (mask[n]) {
      fk = (R)(1e8 * (double)((R)rest[i] - xpos[i]));
    } else {
      fk = f[k] + (R)gforce[i];
      if (a.has_ext) fk += (R)ext[i];
    }
    const R vk = vel[i];
    d.b[3 * (size_t)n + k] = (double)(fk * h + (R)mass[n] * vk);
    d.x[3 * (size_t)n + k] = a.warm ? (double)vk : 0.0;
  }
}

// Per-tet stiffness K = -df/dx (12 x 12, rows 3 a + d, columns 3 b + e over the tet's corners a,
// b and axes d, e) at the positions in planes x: the blocks of the closed form the assembly above
// accumulates, one thread per requested tet (IsotropicMaterial::computeStiffnessMatrixFromPos,
// IsotropicMaterial.cpp:503-552).  ids: solver tet ids; out: 144 FP32 values per requested tet.
template <typename R, int MAT>
__global__ void k_tet_stiffness(DevParams p, int num, const int *__restrict__ ids, const int4 *__restrict__ tets,
                                const float *__restrict__ dmi, const float *__restrict__ vol, const R *__restrict__ x,
                                const float *__restrict__ inst_mu, const float *__restrict__ inst_lam,
                                float *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= num) return;
  const int t = ids[k], N = p.nn, T = p.nt;
  const int4 q4 = tets[t];
  const int qn[4] = {q4.x, q4.y, q4.z, q4.w};
  R Dmi[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) Dmi[i] = (R)dmi[(size_t)i * T + t];
  const R V = (R)vol[t];
  const R mu = (R)inst_mu[0], lam = (R)inst_lam[0];
  R X[4][3];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d) X[c][d] = x[(size_t)d * N + qn[c]];
  R Ds[9], F[9], C[9];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    Ds[3 * d + 0] = X[1][d] - X[0][d];
    Ds[3 * d + 1] = X[2][d] - X[0][d];
    Ds[3 * d + 2] = X[3][d] - X[0][d];
  }
  mul3(Ds, Dmi, F);
  const R J = adj3(F, C);
  const R iJ = (R)1 / J;
  R c[4][3], g[4][3];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    shape_vec(Dmi, m, c[m]);
#pragma unroll
    for (int d = 0; d < 3; ++d) g[m][d] = (c[m][0] * C[d] + c[m][1] * C[3 + d] + c[m][2] * C[6 + d]) * iJ;
  }
  R k1, k2, U[9] = {}, Vv[9] = {}, aw[3] = {};
  const int kk[3] = {0, 0, 1}, ll[3] = {1, 2, 2};
  if (MAT == 0) {
    const R lj = dlog(J);
    k1 = mu - lam * lj;
    k2 = lam;
  } else {
    R S[3];
    svd3(F, U, S, Vv);
    k1 = -lam * (J - (R)1) * J;
    k2 = lam * ((R)2 * J - (R)1) * J;
    const R lim = (sizeof(R) == 4 ? (R)1e-6 : (R)1e-12);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const R den = S[kk[q]] + S[ll[q]];
      aw[q] = (R)1 / (den > lim ? den : lim);
    }
  }
  const R mu2 = (MAT == 0) ? mu : (R)2 * mu;
  // a[m][q][d] = (a^d_kl . c_m) for the pairs kl = 01, 02, 12 (stretch material)
  R a[4][3][3] = {};
  if (MAT != 0)
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          R acc = (R)0;
#pragma unroll
          for (int b = 0; b < 3; ++b)
            acc += (U[3 * d + kk[q]] * Vv[3 * b + ll[q]] - U[3 * d + ll[q]] * Vv[3 * b + kk[q]]) * c[m][b];
          a[m][q][d] = acc;
        }
  float *o = out + 144 * (size_t)k;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const R cc = c[i][0] * c[m][0] + c[i][1] * c[m][1] + c[i][2] * c[m][2];
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int e = 0; e < 3; ++e) {
          R kv = (d == e ? mu2 * cc : (R)0) + k1 * g[m][d] * g[i][e] + k2 * g[i][d] * g[m][e];
          if (MAT != 0) {
            R rot = (R)0;
#pragma unroll
            for (int q = 0; q < 3; ++q) rot += a[i][q][d] * a[m][q][e] * aw[q];
            kv -= mu2 * rot;
          }
          o[12 * (3 * i + d) + 3 * m + e] = (float)(V * kv);
        }
    }
}

// r = b - A x0, z = D^-1 r, p = z; partials r.z, r.r, b.D^-1.b.  The stopping test uses the
// Jacobi-scaled norms |D^-1/2 r| = sqrt(r.z) and |D^-1/2 b|: fixed nodes carry a 1e8 mass and
// spring, and in the plain 2-norm their rows would swamp the free nodes' residual.
template <typename R, int BLK> __global__ void __launch_bounds__(kBlock) k_ie_start(IeDev d) {
  __shared__ double lds[4];
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int n = blk * kBlock + threadIdx.x;
  double rz = 0.0, rr = 0.0, bb = 0.0;
  if (n < d.n) {
    double y[3] = {0.0, 0.0, 0.0};
    const double *__restrict__ xv = d.x;
    row_apply<R, BLK>(d, n, [&](int col, double &x0, double &x1, double &x2) {
      x0 = xv[3 * (size_t)col];
      x1 = xv[3 * (size_t)col + 1];
      x2 = xv[3 * (size_t)col + 2];
    }, y);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const size_t i = 3 * (size_t)n + k;
      const double b = d.b[i];
      const double r = b - y[k];
      const double z = d.dinv[i] * r;
      d.r[i] = r;
      d.z[i] = z;
      d.p[i] = z;
      rz += r * z;
      rr += r * r;
      bb += b * b * d.dinv[i];
    }
  }
  rz = block_sum256(rz, lds);
  rr = block_sum256(rr, lds);
  bb = block_sum256(bb, lds);
  if (threadIdx.x == 0) {
    d.part[4 * (size_t)blk + 0] = rz;
    d.part[4 * (size_t)blk + 1] = rr;
    d.part[4 * (size_t)blk + 2] = bb;
  }
}

__global__ void __launch_bounds__(1024) k_ie_start_ctrl(IeDev d, int nb) {
  __shared__ double lds[48];
  double sums[3];
  partial_sums<3>(d.part, nb, {0, 1, 2}, lds, sums);
  const double rz = sums[0], rr = sums[1], bb = sums[2];
  if (threadIdx.x == 0) {
    IeCtrl c = {};
    c.rz = rz;
    c.rr = rr;
    c.rr0 = rr;
    c.rz0 = rz;
    c.bb = bb;
    c.converged = (rz <= d.tol2 * bb) ? 1 : 0;
    c.done = c.converged || d.maxit <= 0;
    *d.ctrl = c;
  }
}

// Ap = A p, partial p . Ap.
template <typename R, int BLK> __global__ void __launch_bounds__(kBlock) k_ie_spmv(IeDev d) {
  __shared__ double lds[4];
  const IeCtrl *c = d.ctrl;
  if (c->done) return;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int n = blk * kBlock + threadIdx.x;
  double pap = 0.0;
  if (n < d.n) {
    double y[3] = {0.0, 0.0, 0.0};
    const double *__restrict__ p = d.p;
    row_apply<R, BLK>(d, n, [&](int col, double &x0, double &x1, double &x2) {
      x0 = p[3 * (size_t)col];
      x1 = p[3 * (size_t)col + 1];
      x2 = p[3 * (size_t)col + 2];
    }, y);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const size_t i = 3 * (size_t)n + q;
      d.Ap[i] = y[q];
      pap += p[i] * y[q];
    }
  }
  pap = block_sum256(pap, lds);
  if (threadIdx.x == 0) d.part[4 * (size_t)blk + 3] = pap;
}

__global__ void __launch_bounds__(1024) k_ie_alpha(IeDev d, int nb) {
  __shared__ double lds[16];
  if (d.ctrl->done) return;
  double sums[1];
  partial_sums<1>(d.part, nb, {3}, lds, sums);
  const double pap = sums[0];
  if (threadIdx.x == 0) {
    IeCtrl *c = d.ctrl;
    if (!(pap > 0.0)) {  // not positive definite (or a non-finite value): stop without a step
      c->indefinite = 1;
      c->done = 1;
      c->alpha = 0.0;
    } else {
      c->pAp = pap;
      c->alpha = c->rz / pap;
    }
  }
}

// x += alpha p_k, r -= alpha Ap, z = D^-1 r; partials r.z, r.r.  A block covers kUpd = 1024
// consecutive vector entries, 4 per thread 256 apart, so that every load and store of a wavefront
// is one contiguous 512-byte access and there are 4x fewer partial sums to reduce.
constexpr int kUpd = 4 * kBlock;
__global__ void __launch_bounds__(kBlock) k_ie_update(IeDev d) {
  __shared__ double lds[4];
  const IeCtrl *c = d.ctrl;
  if (c->done) return;
  const double alpha = c->alpha;
  const double *__restrict__ pk = d.p;
  const size_t n3 = 3 * (size_t)d.n;
  double rz = 0.0, rr = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const size_t i = (size_t)blockIdx.x * kUpd + j * kBlock + threadIdx.x;
    if (i < n3) {
      d.x[i] += alpha * pk[i];
      const double r = d.r[i] - alpha * d.Ap[i];
      const double z = d.dinv[i] * r;
      d.r[i] = r;
      d.z[i] = z;
      rz += r * z;
      rr += r * r;
    }
  }
  rz = block_sum256(rz, lds);
  rr = block_sum256(rr, lds);
  if (threadIdx.x == 0) {
    d.part[4 * (size_t)blockIdx.x + 0] = rz;
    d.part[4 * (size_t)blockIdx.x + 1] = rr;
  }
}

__global__ void __launch_bounds__(1024) k_ie_beta(IeDev d, int nb) {
  __shared__ double lds[32];
  if (d.ctrl->done) return;
  double sums[2];
  partial_sums<2>(d.part, nb, {0, 1}, lds, sums);
  const double rz = sums[0], rr = sums[1];
  if (threadIdx.x == 0) {
    IeCtrl *c = d.ctrl;
    c->k += 1;
    c->rr = rr;
    c->beta = rz / c->rz;
    c->rz = rz;
    c->converged = (rz <= d.tol2 * c->bb) ? 1 : 0;
    c->done = c->converged || c->k >= d.maxit;
  }
}

// p = z + beta p for the next iteration (after k_ie_beta), one thread per vector entry.
__global__ void __launch_bounds__(kBlock) k_ie_pupdate(IeDev d) {
  const IeCtrl *c = d.ctrl;
  if (c->done) return;
  const double beta = c->beta;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * (size_t)d.n) d.p[i] = d.z[i] + beta * d.p[i];
}

// v = solution, x += h v for every node (cublasSaxpy in doBackEuler, CUDAInterface.cu:873).
template <typename R>
__global__ void k_ie_finish(DevParams p, float h, IeDev d, R *__restrict__ xpos, R *__restrict__ vel) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= d.n) return;
  const bool zero = d.ctrl->bb == 0.0;  // b = 0: the solution is 0 whatever the start
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const size_t i = (size_t)k * p.nn + n;
    const R v = zero ? (R)0 : (R)d.x[3 * (size_t)n + k];
    vel[i] = v;
    xpos[i] = xpos[i] + (R)h * v;
  }
}

// ---------------------------------------------------------------------------------------------
// Projective dynamics (ProjDynamic.cpp).  The global matrix (buildGlobalSolverMatrix, :46-96) is
// G = M / t^2 + sum_k w_k Ac_k^T Ac_k with w_k = stiffness * vol_k; Ac_k^T Ac_k couples corners i, j
// of tet k by (c_i . c_j) times the 3x3 identity, so G is stored with one value per slot.
// ---------------------------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_pd_matrix(DevParams p, float stiffness, IeDev d, const int4 *__restrict__ tets, const float *__restrict__ dmi,
                const float *__restrict__ vol, const int *__restrict__ inc_off, const int *__restrict__ inc,
                const float *__restrict__ mass) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= d.n) return;
  R *__restrict__ A = (R *)d.A;
  const int T = p.nt;
  const int lane = n & 63;
  const int g0 = d.slice_off[n >> 6];
  const int width = d.slice_off[(n >> 6) + 1] - g0;
  const int len = d.row_len[n];
  for (int j = 0; j < width; ++j) A[(size_t)(g0 + j) * 64 + lane] = (R)0;
  for (int e = inc_off[n]; e < inc_off[n + 1]; ++e) {
    const int t = inc[e] >> 2, ci = inc[e] & 3;
    const int4 q4 = tets[t];
    const int qn[4] = {q4.x, q4.y, q4.z, q4.w};
    R Dmi[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Dmi[k] = (R)dmi[(size_t)k * T + t];
    const R w = (R)stiffness * (R)vol[t];
    R c[4][3];
#pragma unroll
    for (int m = 0; m < 4; ++m) shape_vec(Dmi, m, c[m]);
    R cs[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) cs[k] = ci == 0 ? c[0][k] : ci == 1 ? c[1][k] : ci == 2 ? c[2][k] : c[3][k];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int slot_g = g0 + find_col(d.bcol, g0, lane, len, qn[m]);
      A[(size_t)slot_g * 64 + lane] += w * (cs[0] * c[m][0] + cs[1] * c[m][1] + cs[2] * c[m][2]);
    }
  }
  const R h = (R)p.h;
  const R t2inv = (R)1 / (h * h);
  const int ds = d.diag_slot[n];
  A[ds] += (R)mass[n] * t2inv;
  const double di = 1.0 / (double)A[ds];
#pragma unroll
  for (int k = 0; k < 3; ++k) d.dinv[3 * (size_t)n + k] = di;
}

// Momentum term of doProjDynamics (ProjDynamic.cpp:442-460): s = (x + t v + t^2 m^-1 f_ext) m / t^2
// with f_ext = gravity + external forces (TetMesh::computeExternalForces; fixed nodes have
// m^-1 = 0); the iterate starts at x.
template <typename R>
__global__ void k_pd_begin(DevParams p, IeDev d, double *__restrict__ srhs, const R *__restrict__ xpos,
                           const R *__restrict__ vel, const float *__restrict__ mass, const float *__restrict__ gforce,
                           const float *__restrict__ ext, int has_ext, const int *__restrict__ mask) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= d.n) return;
  const R t = (R)p.h, m = (R)mass[n];
  const R im = mask[n] ? (R)0 : (R)1 / m;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const size_t i = (size_t)k * p.nn + n;
    R f = (R)gforce[i];
    if (has_ext) f += (R)ext[i];
    const R y = xpos[i] + t * vel[i] + t * t * im * f;
    srhs[3 * (size_t)n + k] = (double)(y * m / (t * t));
    d.x[3 * (size_t)n + k] = (double)xpos[i];
  }
}

// The projection rotation of one tet: R = U V^T of F's SVD with the signs the reference fixes
// (ProjDynamic.cpp:359-368: a column of U or V whose determinant is negative is negated, column 0
// of the singular values sorted in decreasing order).  With svd3's rotations U, V and the sign of
// det F on the last singular value this is U V^T for det F >= 0 and U diag(-1, 1, -1) V^T otherwise.
template <typename R> __device__ __forceinline__ void pd_rotation(const R *F, R *Rot) {
  // det F > 0: U V^T is the rotation of the polar decomposition (scaled Newton iteration)
  if (polar_rotation(F, Rot)) return;
  R U[9], S[3], V[9];
  svd3(F, U, S, V);
  const R s0 = S[2] < (R)0 ? (R)-1 : (R)1;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Rot[3 * i + j] = s0 * U[3 * i] * V[3 * j] + U[3 * i + 1] * V[3 * j + 1] + s0 * U[3 * i + 2] * V[3 * j + 2];
}

// Local step (projectLocalConstraints2, ProjDynamic.cpp:223-308), in two passes like the
// reference's: k_pd_tet evaluates each tet once -- F from the iterate pos (3N node-major
// doubles), the SVD, R -- and writes its four corner projections w_k R_k c_i to rec (12 values
// per tet); k_pd_gather sums each node's corner records over its tets in ascending order,
// out = s + sum (s may be null: the projection alone).
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_pd_tet(DevParams p, float stiffness, const double *__restrict__ pos, const int4 *__restrict__ tets,
             const float *__restrict__ dmi, const float *__restrict__ vol, const int *__restrict__ inc_pos,
             R *__restrict__ rec) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int T = p.nt;
  if (t >= T) return;
  const int4 q4 = tets[t];
  const int qn[4] = {q4.x, q4.y, q4.z, q4.w};
  R X[4][3];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int k = 0; k < 3; ++k) X[c][k] = (R)pos[3 * (size_t)qn[c] + k];
  R Dmi[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) Dmi[k] = (R)dmi[(size_t)k * T + t];
  R Ds[9], F[9], Rot[9];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    Ds[3 * k + 0] = X[1][k] - X[0][k];
    Ds[3 * k + 1] = X[2][k] - X[0][k];
    Ds[3 * k + 2] = X[3][k] - X[0][k];
  }
  mul3(Ds, Dmi, F);
  pd_rotation(F, Rot);
  const R w = (R)stiffness * (R)vol[t];
#pragma unroll
  for (int ci = 0; ci < 4; ++ci) {
    R c[3];
    shape_vec(Dmi, ci, c);
    R *o = rec + 3 * (size_t)inc_pos[4 * (size_t)t + ci];  // the (tet, corner) slot of its node's run
#pragma unroll
    for (int k = 0; k < 3; ++k) o[k] = w * (Rot[3 * k] * c[0] + Rot[3 * k + 1] * c[1] + Rot[3 * k + 2] * c[2]);
  }
}

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_pd_gather(int n_nodes, const int *__restrict__ inc_off, const R *__restrict__ rec,
                const double *__restrict__ srhs, double *__restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_nodes) return;
  R acc[3] = {(R)0, (R)0, (R)0};
  for (int e = inc_off[n]; e < inc_off[n + 1]; ++e) {  // the node's run, in ascending tet order
    const R *r = rec + 3 * (size_t)e;
#pragma unroll
    for (int k = 0; k < 3; ++k) acc[k] += r[k];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const size_t i = 3 * (size_t)n + k;
    out[i] = srhs ? (double)((R)srhs[i] + acc[k]) : (double)acc[k];
  }
}

// v = damping (x_new - x) / t, x = x_new (ProjDynamic.cpp:477-478), every node.
template <typename R>
__global__ void k_pd_finish(DevParams p, float damping, IeDev d, R *__restrict__ xpos, R *__restrict__ vel) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= d.n) return;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const size_t i = (size_t)k * p.nn + n;
    const R xn = (R)d.x[3 * (size_t)n + k];
    vel[i] = (xn - xpos[i]) / (R)p.h * (R)damping;
    xpos[i] = xn;
  }
}

inline int nb_of(int n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

template <typename R>
void IeLaunch<R>::tet_stiffness(const DevMesh<R> &m, const DevParams &p, const float *inst_mu, const float *inst_lam,
                                const R *x, int num, const int *ids, float *out, hipStream_t st) {
  if (num <= 0) return;
  const int blocks = (num + 63) / 64;
  if (p.material == 0)
    hipLaunchKernelGGL((k_tet_stiffness<R, 0>), dim3(blocks), dim3(64), 0, st, p, num, ids, m.tets, m.dmi, m.vol, x,
                       inst_mu, inst_lam, out);
  else
    hipLaunchKernelGGL((k_tet_stiffness<R, 1>), dim3(blocks), dim3(64), 0, st, p, num, ids, m.tets, m.dmi, m.vol, x,
                       inst_mu, inst_lam, out);
}
template <typename R>
void IeLaunch<R>::assemble(const DevMesh<R> &m, const TileDev &td, const DevParams &p, const IeDev &d,
                           const IeAsm &a, int max_width, hipStream_t st) {
  // rows accumulated in LDS when the widest slice fits in 96 KB per workgroup, else in place
  // (6.3M-tet bar, 15 slots per row: FP32 5.8 -> 3.1 ms, FP64 8.5 -> 4.6 ms)
  const size_t lds = (size_t)max_width * 9 * 64 * sizeof(R);
  if (lds <= 98304) {
    const dim3 grid((d.n + 63) / 64), blk(64);
    if (p.material == 0)
      hipLaunchKernelGGL((k_ie_assemble<R, 0, true>), grid, blk, lds, st, p, a, d, m.tets, m.dmi, m.vol, m.inc_off,
                         m.inc, m.xpos, m.v, m.rest, m.mass, m.gforce, m.ext, m.mask, td.inst_mu, td.inst_lam);
    else
      hipLaunchKernelGGL((k_ie_assemble<R, 1, true>), grid, blk, lds, st, p, a, d, m.tets, m.dmi, m.vol, m.inc_off,
                         m.inc, m.xpos, m.v, m.rest, m.mass, m.gforce, m.ext, m.mask, td.inst_mu, td.inst_lam);
    return;
  }
  if (p.material == 0)
    hipLaunchKernelGGL((k_ie_assemble<R, 0, false>), dim3(nb_of(d.n)), dim3(kBlock), 0, st, p, a, d, m.tets, m.dmi,
                       m.vol, m.inc_off, m.inc, m.xpos, m.v, m.rest, m.mass, m.gforce, m.ext, m.mask, td.inst_mu,
                       td.inst_lam);
  else
    hipLaunchKernelGGL((k_ie_assemble<R, 1, false>), dim3(nb_of(d.n)), dim3(kBlock), 0, st, p, a, d, m.tets, m.dmi,
                       m.vol, m.inc_off, m.inc, m.xpos, m.v, m.rest, m.mass, m.gforce, m.ext, m.mask, td.inst_mu,
                       td.inst_lam);
}

template <typename R> void IeLaunch<R>::start(const IeDev &d, hipStream_t st) {
  if (d.scalar) hipLaunchKernelGGL((k_ie_start<R, 1>), dim3(nb_of(d.n)), dim3(kBlock), 0, st, d);
  else hipLaunchKernelGGL((k_ie_start<R, 3>), dim3(nb_of(d.n)), dim3(kBlock), 0, st, d);
  hipLaunchKernelGGL(k_ie_start_ctrl, dim3(1), dim3(1024), 0, st, d, nb_of(d.n));
}

// e0 set: the launch records the dispatch's start and end in e0, e1 (as a kernel trace times it)
template <typename F>
static void ie_launch(F kernel, dim3 grid, dim3 block, hipStream_t st, hipEvent_t e0, hipEvent_t e1, const IeDev &d) {
  if (e0) hipExtLaunchKernelGGL(kernel, grid, block, 0u, st, e0, e1, 0u, d);
  else hipLaunchKernelGGL(kernel, grid, block, 0, st, d);
}
template <typename R> void IeLaunch<R>::spmv(const IeDev &d, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
  if (d.scalar) ie_launch(k_ie_spmv<R, 1>, dim3(nb_of(d.n)), dim3(kBlock), st, e0, e1, d);
  else ie_launch(k_ie_spmv<R, 3>, dim3(nb_of(d.n)), dim3(kBlock), st, e0, e1, d);
}
template <typename R> void IeLaunch<R>::alpha(const IeDev &d, hipStream_t st) {
  hipLaunchKernelGGL(k_ie_alpha, dim3(1), dim3(1024), 0, st, d, nb_of(d.n));
}
template <typename R> void IeLaunch<R>::update(const IeDev &d, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
  ie_launch(k_ie_update, dim3((3 * d.n + kUpd - 1) / kUpd), dim3(kBlock), st, e0, e1, d);
}
template <typename R> void IeLaunch<R>::beta(const IeDev &d, hipStream_t st) {
  hipLaunchKernelGGL(k_ie_beta, dim3(1), dim3(1024), 0, st, d, (3 * d.n + kUpd - 1) / kUpd);
  hipLaunchKernelGGL(k_ie_pupdate, dim3(nb_of(3 * d.n)), dim3(kBlock), 0, st, d);
}

template <typename R>
void IeLaunch<R>::finish(const DevMesh<R> &m, const DevParams &p, float h, const IeDev &d, hipStream_t st) {
  hipLaunchKernelGGL(k_ie_finish<R>, dim3(nb_of(d.n)), dim3(kBlock), 0, st, p, h, d, m.xpos, m.v);
}

template <typename R>
void IeLaunch<R>::pd_matrix(const DevMesh<R> &m, const DevParams &p, float stiffness, const IeDev &d, hipStream_t st) {
  hipLaunchKernelGGL(k_pd_matrix<R>, dim3(nb_of(d.n)), dim3(kBlock), 0, st, p, stiffness, d, m.tets, m.dmi, m.vol,
                     m.inc_off, m.inc, m.mass);
}
template <typename R>
void IeLaunch<R>::pd_begin(const DevMesh<R> &m, const DevParams &p, const IeDev &d, double *srhs, int has_ext,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_pd_begin<R>, dim3(nb_of(d.n)), dim3(kBlock), 0, st, p, d, srhs, m.xpos, m.v, m.mass, m.gforce,
                     m.ext, has_ext, m.mask);
}
template <typename R>
void IeLaunch<R>::pd_local(const DevMesh<R> &m, const DevParams &p, float stiffness, const int *inc_pos,
                           const double *pos, const double *srhs, double *out, hipStream_t st) {
  // the per-tet force records of the standalone calls (12 R per tet) serve as the staging array,
  // filled in incidence order
  hipLaunchKernelGGL(k_pd_tet<R>, dim3(nb_of(p.nt)), dim3(kBlock), 0, st, p, stiffness, pos, m.tets, m.dmi, m.vol,
                     inc_pos, m.frec);
  hipLaunchKernelGGL(k_pd_gather<R>, dim3(nb_of(p.nn)), dim3(kBlock), 0, st, p.nn, m.inc_off, (const R *)m.frec, srhs,
                     out);
}
template <typename R>
void IeLaunch<R>::pd_finish(const DevMesh<R> &m, const DevParams &p, float damping, const IeDev &d, hipStream_t st) {
  hipLaunchKernelGGL(k_pd_finish<R>, dim3(nb_of(d.n)), dim3(kBlock), 0, st, p, damping, d, m.xpos, m.v);
}

template struct IeLaunch<float>;
template struct IeLaunch<double>;

}  // namespace pbs
