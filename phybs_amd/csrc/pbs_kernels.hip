// pbs_kernels.hip -- HIP kernels of the descent solver for gfx950 (CDNA4, wave64).
//
// One descent pass = k_tile (each tet once, per tile: force, energy, Hessian diagonal on refresh
// passes, gathered per node in LDS) -> k_node (owned + halo partials, Jacobi step, Chebyshev,
// energy and |g|^2 partials) -> k_control (FP64 reduction in a fixed order and the line-search /
// Chebyshev / buffer-role decisions).  This replaces the reference's computeForceKd /
// updateNodesNextIter / updateChebyshev / thrust::reduce and the host loop of
// CUDAInterface::doDescentOpt (CUDAInterface.cu:482-1007) with no host round trip and no atomics.
// k_tet + k_gather evaluate forces and the Hessian diagonal for the standalone material calls.
//
// Layout: nodes are renumbered in Morton order and tets sorted by their smallest node; node
// vectors are 3 planes (x | y | z) of nn values (SoA); per-tet rest data are 10 planes (Dm^-1
// entries, volume) plus 16 bytes per tet of tile-local ids and contribution slots for the tile
// pass; per-tet force records of the standalone calls are 12 consecutive values.
#include <algorithm>
#include <type_traits>

#include <hip/hip_ext.h>

#include "pbs_device.h"
#include "pbs_internal.h"

namespace pbs {

// ---------------------------------------------------------------------------------------------
// Deterministic block reduction of a double: wave64 shuffle tree, then the waves' sums in order.
// The result is valid in thread 0.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double block_sum(double v, double *lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) r += lds[i];
  }
  __syncthreads();
  return r;
}

// N block sums at once (one LDS exchange): the same order per value as block_sum; lds holds 16 N.
template <int N> __device__ __forceinline__ void block_sums(double (&v)[N], double *lds) {
#pragma unroll
  for (int j = 0; j < N; ++j)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[j] += __shfl_down(v[j], off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < N; ++j) lds[16 * j + w] = v[j];
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double r = 0.0;
      for (int i = 0; i < nw; ++i) r += lds[16 * j + i];
      v[j] = r;
    }
  }
  __syncthreads();
}

// Chebyshev schedule of the reference (CUDAInterface.cu:967-986; DescentOptimize.cpp:156-175).
__device__ __forceinline__ void cheb_schedule(const DevParams &p, int k, float &rho, float &omega) {
  if (k == 0) {
    rho = 0.0f;
    omega = 1.0f;
  }
  omega = 4.0f / (4.0f - rho * rho * omega);
  for (int i = 0; i < p.profile_n; ++i) {
    if (k == p.profile_k[i] - 1) {
      rho = 0.0f;
      omega = 1.0f;
    }
    if (k == p.profile_k[i]) {
      rho = p.profile_rho[i];
      omega = 2.0f / (2.0f - rho * rho);
      break;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Setup: rest-shape data (compute_DmInv_ANs_Mass, CUDAInterface.cu:273-325; TetMesh.cpp:189-216).
// Contraction is off so that Dm^-1, volumes and masses round exactly as the FP32 reference
// formulas written out without fused multiply-adds.
// ---------------------------------------------------------------------------------------------
#pragma clang fp contract(off)
// Dm^-1 (row-major) and the volume |det Dm| / 6 of a rest tet from its corners' rest positions
// r[c][k] (compute_DmInv_ANs_Mass, CUDAInterface.cu:273-325; TetMesh.cpp:189-216); returns det Dm.
// The setup kernel and the tile pass both call this, so the tile pass's values are bitwise the
// stored ones.
__device__ __forceinline__ float rest_shape(const float r[4][3], float Dmi[9], float &V) {
  float Dm[9], C[9];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float x0 = r[0][k];
    Dm[3 * k + 0] = r[1][k] - x0;
    Dm[3 * k + 1] = r[2][k] - x0;
    Dm[3 * k + 2] = r[3][k] - x0;
  }
  // adjugate and determinant written out here so that the no-contraction pragma applies
  C[0] = Dm[4] * Dm[8] - Dm[7] * Dm[5];
  C[1] = Dm[7] * Dm[2] - Dm[1] * Dm[8];
  C[2] = Dm[1] * Dm[5] - Dm[4] * Dm[2];
  C[3] = Dm[5] * Dm[6] - Dm[3] * Dm[8];
  C[4] = Dm[0] * Dm[8] - Dm[2] * Dm[6];
  C[5] = Dm[2] * Dm[3] - Dm[0] * Dm[5];
  C[6] = Dm[3] * Dm[7] - Dm[4] * Dm[6];
  C[7] = Dm[1] * Dm[6] - Dm[0] * Dm[7];
  C[8] = Dm[0] * Dm[4] - Dm[1] * Dm[3];
  const float det = Dm[0] * C[0] + Dm[3] * C[1] + Dm[6] * C[2];
  V = fabsf(det) / 6.0f;
  const float inv = 1.0f / det;
#pragma unroll
  for (int i = 0; i < 9; ++i) Dmi[i] = C[i] * inv;
  return det;
}

__global__ void k_setup_tets(DevParams p, const float *__restrict__ rest, const int4 *__restrict__ tets,
                             float *__restrict__ dmi, float *__restrict__ vol, int *__restrict__ flag) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= p.nt) return;
  const int N = p.nn, T = p.nt;
  const int4 q = tets[t];
  if (q.x < 0 || q.y < 0 || q.z < 0 || q.w < 0 || q.x >= N || q.y >= N || q.z >= N || q.w >= N) {
    atomicOr(flag, 1);
    return;
  }
  const int qn[4] = {q.x, q.y, q.z, q.w};
  float r[4][3];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int k = 0; k < 3; ++k) r[c][k] = rest[k * N + qn[c]];
  float Dmi[9], V;
  const float det = rest_shape(r, Dmi, V);
  if (!(det > 0.0f)) atomicOr(flag, 2);
  vol[t] = V;
#pragma unroll
  for (int i = 0; i < 9; ++i) dmi[(size_t)i * T + t] = Dmi[i];
}

// Lumped mass gathered over the incidence list in ascending tet order (the reference's float
// atomics, CUDAInterface.cu:297-301, sum in an unspecified order); gravity force from the mass
// before constrained nodes get fixed_mass (CUDAInterface.cu:716-719).
__global__ void k_setup_nodes(DevParams p, const int *__restrict__ inc_off, const int *__restrict__ inc,
                              const float *__restrict__ vol, const int *__restrict__ mask, float *__restrict__ mass,
                              float *__restrict__ mnode, float *__restrict__ gforce) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= p.nn) return;
  float m = 0.0f;
  for (int e = inc_off[n]; e < inc_off[n + 1]; ++e) {
    const float w = vol[inc[e] >> 2] / 4.0f * p.density;
    m += w;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) gforce[(size_t)d * p.nn + n] = p.grav[d] * m;
  mass[n] = mask[n] ? p.fixed_mass : m;
  mnode[n] = mask[n] ? -1.0f : m;
}
#pragma clang fp contract(on)

// ---------------------------------------------------------------------------------------------
// Per-tet pass: forces, energy, optional Hessian diagonal (computeForceKd, CUDAInterface.cu:495-568).
// One thread per tet; node positions are gathered from the 3 planes, rest data read coalesced.
// ---------------------------------------------------------------------------------------------
template <typename R, int MAT>
__global__ void __launch_bounds__(256) k_tet(DevParams p, const Ctrl *__restrict__ ctrl, Ring<R> ring,
                                             const int4 *__restrict__ tets, const float *__restrict__ dmi,
                                             const float *__restrict__ vol, R *__restrict__ frec,
                                             float *__restrict__ hrec, double *__restrict__ etet,
                                             R *__restrict__ tete, int force_hd, TileDev td) {
  __shared__ double lds[16];
  if (ctrl->done) return;
  const int hd = force_hd >= 0 ? force_hd : ctrl->hd_now;
  const R *__restrict__ x = ring.b[ctrl->cur];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = p.nn, T = p.nt;
  double e = 0.0;
  if (t < T) {
    const int4 q = tets[t];
    R Dmi[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Dmi[i] = (R)dmi[(size_t)i * T + t];
    const R V = (R)vol[t];
    R Ds[9];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const R *xk = x + (size_t)k * N;
      const R x0 = xk[q.x];
      Ds[3 * k + 0] = xk[q.y] - x0;
      Ds[3 * k + 1] = xk[q.z] - x0;
      Ds[3 * k + 2] = xk[q.w] - x0;
    }
    R F[9];
    mul3(Ds, Dmi, F);
    R f12[12], h12[12], en;
    int lo = 0, hi = td.ninst;  // instance of tet t
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (td.inst_tet_off[mid] <= t) lo = mid;
      else hi = mid;
    }
    const R mu = (R)td.inst_mu[lo], lam = (R)td.inst_lam[lo];
    if (hd) {
      if (MAT == 0) neo_tet<R, true>(F, Dmi, V, mu, lam, f12, en, h12);
      else stretch_tet<R, true>(F, Dmi, V, mu, lam, f12, en, h12);
      float *hr = hrec + 12 * (size_t)t;
#pragma unroll
      for (int i = 0; i < 12; ++i) hr[i] = (float)h12[i];
    } else {
      if (MAT == 0) neo_tet<R, false>(F, Dmi, V, mu, lam, f12, en, h12);
      else stretch_tet<R, false>(F, Dmi, V, mu, lam, f12, en, h12);
    }
    R *fr = frec + 12 * (size_t)t;
#pragma unroll
    for (int i = 0; i < 12; ++i) fr[i] = f12[i];
    if (p.clamp == 1) en = en > (R)0 ? en : (R)0;                  // CUDAInterface.cu:549
    else if (p.clamp == 2) en = en > (R)1e-6 ? en : (R)0;          // IsotropicNeohookeanMaterial.cpp:129
    if (tete) tete[t] = en;
    e = (double)en;
  }
  const double s = block_sum(e, lds);
  if (threadIdx.x == 0) etet[blockIdx.x] = s;
}

// ---------------------------------------------------------------------------------------------
// Loop control of the reference (CUDAInterface.cu:917-1003) applied on the device after the
// pass over iterate x_k: E = incremental potential at x_k (FP64), g2 = |g(x_k)|^2.
// ---------------------------------------------------------------------------------------------
__device__ void control_apply(const DevParams &p, Ctrl &C, double E, double g2) {
  const double g = sqrt(g2);
  C.g = g;
  if (C.first) {
    if (C.g0 < 0.0) C.g0 = g;
    C.first = 0;
  }
  // (mixed mode: convergence is only declared on a pass with FP64 forces)
  if (p.relgrad_tol > 0.0 && !p.f32math && g <= p.relgrad_tol * C.g0) {
    C.converged = 1;
    C.done = 1;
    C.passes++;
    return;
  }
  int rb = 0;
  const bool robust = p.relgrad_tol > 0.0;
  if (C.check_now) {
    C.E = E;
    // Rel-grad mode (no reference counterpart) also rejects a non-finite energy.
    const bool worse = robust ? !(E <= C.E_old) : (E > C.E_old);
    if (C.k != 0 && worse) {  // rollback (CUDAInterface.cu:946-958)
      C.alpha *= p.alpha_decay;
      if (robust) C.kc = C.k;  // rel-grad mode: keep the budget, restart Chebyshev at k
      else C.budget -= C.k - p.energy_check;
      C.cur = C.old;
      C.rollbacks++;
      rb = 1;
    } else {
      C.E_old = E;
      C.old = C.cur;
    }
  }
  if (!rb) {
    C.rho = C.rho_next;
    C.omega = C.omega_next;
    C.last = C.cur;
    C.cur = C.nxt;
    int f = 0;
    while (f == C.cur || f == C.last || f == C.old) ++f;
    C.nxt = f;
    C.k++;
  }
  C.passes++;
  if (C.alpha < p.alpha_min || !(C.k < C.budget)) {
    C.done = 1;
  } else {
    C.hd_now = (C.k % p.hd_refresh) == 0;
    C.check_now = (C.k % p.energy_check) == 0;
    float rho = C.rho, om = C.omega;
    cheb_schedule(p, C.k - C.kc, rho, om);
    C.rho_next = rho;
    C.omega_next = om;
    C.cheb_now = om != 1.0f;
  }
}

// The fields of a control block that a descent pass's tile and node kernels read.
struct PassCtl {
  int done, cur, nxt, last, hd_now, check_now, cheb_now;
  float alpha, omega_next;
};
__device__ __forceinline__ PassCtl pass_ctl(const Ctrl *c) {
  return {c->done, c->cur, c->nxt, c->last, c->hd_now, c->check_now, c->cheb_now, c->alpha, c->omega_next};
}

// XCD-aware tile order: workgroups are dispatched round-robin over the 8 XCDs, so block i runs
// on XCD i % 8.  Giving each XCD a contiguous range of tiles keeps neighbouring tiles (which
// share halo nodes and halo tets) in the same 4 MB L2.
__device__ __forceinline__ int xcd_tile(int bid, int nblk) {
  const int per = (nblk + 7) >> 3;
  const int xcd = bid & 7, k = bid >> 3;
  const int full = nblk - 8 * (per - 1);  // XCDs that get `per` tiles (the rest get per - 1)
  const int start = xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1);
  return start + k;
}

// One tet of a tile: F from the tile's LDS positions (stride 4 R per local node), the material
// response, and its four corner contributions written to their slots of the LDS array rec (W
// values each): forces on phase 0 (plus the clamped energy on check passes), Hessian-diagonal
// entries rounded to FP32 on phase 1.  The arithmetic type C is R, or FP32 on FP64 positions (the
// mixed mode).  Mixed mode: F is formed in R; the energy of check passes is evaluated in R from
// that F (the same value the FP64 passes compute, so the line search compares like with like);
// the neo-Hookean force is evaluated in FP32 from the displacement gradient H = F - I, whose FP32
// rounding is relative to the strain instead of to 1; contributions are stored as FP32.
template <typename R, typename C, int MAT>
__device__ __forceinline__ void tile_tet(int phase, const uint4 lp, const float Dmf[9], float Vf,
                                         const R *__restrict__ sx, C *__restrict__ rec, C mu, C lam, int chk, int clamp,
                                         double &e_el) {
  constexpr int W = ContribWidth<C>::value;
  constexpr bool MIX = sizeof(C) < sizeof(R);
  const R *p0 = sx + 4 * (lp.x & 0xffff), *p1 = sx + 4 * (lp.x >> 16), *p2 = sx + 4 * (lp.y & 0xffff),
          *p3 = sx + 4 * (lp.y >> 16);
  typedef typename std::conditional<MIX, R, C>::type FT;  // type F is formed in
  FT Ds[9], Dm[9], Fw[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) Dm[k] = (FT)Dmf[k];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const R x0 = p0[k];
    Ds[3 * k + 0] = (FT)(p1[k] - x0);
    Ds[3 * k + 1] = (FT)(p2[k] - x0);
    Ds[3 * k + 2] = (FT)(p3[k] - x0);
  }
  mul3(Ds, Dm, Fw);
  C F[9], Dmi[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    Dmi[k] = (C)Dmf[k];
    F[k] = (C)Fw[k];
  }
  const C V = (C)Vf;
  C f12[12], h12[12], en = (C)0;
  const unsigned q[4] = {lp.z & 0xffff, lp.z >> 16, lp.w & 0xffff, lp.w >> 16};
  if (phase == 0) {
    if (MIX && MAT == 0) {
      C H[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) H[k] = (C)(Fw[k] - ((k & 3) == 0 ? (FT)1 : (FT)0));
      neo_tet_disp<C>(H, Dmi, V, mu, lam, f12);
    } else if (MAT == 0) {
      neo_tet<C, false>(F, Dmi, V, mu, lam, f12, en, h12);
    } else {
      stretch_tet<C, false>(F, Dmi, V, mu, lam, f12, en, h12);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      C *r = rec + W * q[c];
      r[0] = f12[3 * c];
      r[1] = f12[3 * c + 1];
      r[2] = f12[3 * c + 2];
    }
    if (chk) {
      R er = MIX ? tet_energy<R, MAT>((const R *)Fw, (R)Vf, (R)mu, (R)lam) : (R)en;
      if (clamp == 1) er = er > (R)0 ? er : (R)0;          // CUDAInterface.cu:549
      else if (clamp == 2) er = er > (R)1e-6 ? er : (R)0;  // IsotropicNeohookeanMaterial.cpp:129
      e_el += (double)er;
    }
  } else if (MAT == 0) {
    // the neo-Hookean stiffness diagonal of neo_tet<C, true>, written corner by corner as it is
    // formed (the 12 values are never all live: this phase sets the kernel's register count)
    C Cf[9];
    const C J = adj3(F, Cf);
    const C iJ = (C)1 / J;
    C Fi[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Fi[i] = Cf[i] * iJ;
    const C k2 = mu + lam - lam * dlog(J);
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      C c3[3];
      if (n == 0) {
#pragma unroll
        for (int j = 0; j < 3; ++j) c3[j] = -((C)Dmf[j] + (C)Dmf[3 + j] + (C)Dmf[6 + j]);
      } else {
#pragma unroll
        for (int j = 0; j < 3; ++j) c3[j] = (C)Dmf[3 * (n - 1) + j];
      }
      const C cc = c3[0] * c3[0] + c3[1] * c3[1] + c3[2] * c3[2];
      C h[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const C sd = c3[0] * Fi[d] + c3[1] * Fi[3 + d] + c3[2] * Fi[6 + d];
        h[d] = V * (mu * cc + k2 * sd * sd);
      }
      C *r = rec + W * q[n];
      r[0] = (C)(float)h[0];
      r[1] = (C)(float)h[1];
      r[2] = (C)(float)h[2];
      // FP64: the store stays here (sunk to the end of the loop it would hold all 12 values)
      if constexpr (sizeof(C) == 8) __builtin_amdgcn_sched_barrier(0);
    }
  } else {
    stretch_tet<C, true>(F, Dmi, V, mu, lam, f12, en, h12);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      C *r = rec + W * q[c];
      r[0] = (C)(float)h12[3 * c];
      r[1] = (C)(float)h12[3 * c + 1];
      r[2] = (C)(float)h12[3 * c + 2];
    }
  }
}

// Gather of one tile: each thread sums the contribution runs of its local nodes in order and
// writes one partial per owned node (3 planes of N) and per halo node (xyz at its slot); the sums
// are formed in R whatever the contributions' type Q.
template <typename R, typename Q>
__device__ __forceinline__ void tile_gather(int phase, const TileInfo &ti, const TileDev &td, const Q *__restrict__ rec,
                                            int N, R *__restrict__ fown, R *__restrict__ hpart,
                                            float *__restrict__ fownh, float *__restrict__ hparth) {
  constexpr int W = ContribWidth<Q>::value;
  const int L = ti.own_cnt + ti.halo_cnt;
  const unsigned short *off = td.tile_off + ti.off_start;
  for (int i = threadIdx.x; i < L; i += blockDim.x) {
    const int o0 = off[i], o1 = off[i + 1];
    size_t d0, stride;
    if (i < ti.own_cnt) {
      d0 = (size_t)ti.own_start + i;
      stride = N;
    } else {
      d0 = 3 * (size_t)td.halo_slot[ti.halo_off + (i - ti.own_cnt)];
      stride = 1;
    }
    // the run is summed in order; FP32 contributions are read 4 at a time ahead of their additions
    // (slots past the run clamped to its first slot and not added: FP32 tile pass 85.6 -> 82.4 us
    // at 6.3 M tets), FP64 ones one at a time (4 ahead: 148 -> 152 us)
    constexpr int U = sizeof(Q) == 4 ? 4 : 1;
    if (phase == 0) {
      R a0 = (R)0, a1 = (R)0, a2 = (R)0;
      for (int j = o0; j < o1; j += U) {
        Q t[U][3];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const Q *r = rec + W * (j + u < o1 ? j + u : o0);
          t[u][0] = r[0];
          t[u][1] = r[1];
          t[u][2] = r[2];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (j + u < o1) {
            a0 += (R)t[u][0];
            a1 += (R)t[u][1];
            a2 += (R)t[u][2];
          }
      }
      R *o = i < ti.own_cnt ? fown : hpart;
      o[d0] = a0;
      o[stride + d0] = a1;
      o[2 * stride + d0] = a2;
    } else {
      float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
      for (int j = o0; j < o1; j += U) {
        Q t[U][3];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const Q *r = rec + W * (j + u < o1 ? j + u : o0);
          t[u][0] = r[0];
          t[u][1] = r[1];
          t[u][2] = r[2];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (j + u < o1) {
            a0 += (float)t[u][0];
            a1 += (float)t[u][1];
            a2 += (float)t[u][2];
          }
      }
      float *o = i < ti.own_cnt ? fownh : hparth;
      o[d0] = a0;
      o[stride + d0] = a1;
      o[2 * stride + d0] = a2;
    }
  }
}

// Byte offset of the tile pass's rest-position region in LDS (after the positions and the
// contribution slots; 16-byte aligned).
template <typename R> __host__ __device__ inline size_t tile_rest_offset(const TileDev &td) {
  const size_t b = (4 * (size_t)td.lcap + 4 * (size_t)ContribWidth<R>::value * td.tcap) * sizeof(R);
  return (b + 15) & ~(size_t)15;
}

// ---------------------------------------------------------------------------------------------
// Tile pass (per-tet half of computeForceKd, CUDAInterface.cu:495-568): one workgroup per tile
// evaluates each of the tile's tets once -- force, energy on check passes, Hessian diagonal on
// refresh passes -- and writes every corner contribution into its precomputed slot of an LDS
// array in which each local node's contributions form one contiguous run (ordered by tet, then
// corner).  The gather then sums each run in that fixed order (no atomics) and writes one partial
// per owned node (fown) and per halo node (hpart).
//   LDS: positions of local nodes (4 R each), contributions (4 per tet, W values each), rest
//   positions of local nodes (float4 each).
//   HBM per tet: 16 B of local ids / slots.  Dm^-1 and the volume are not streamed from memory
//   (40 B per tet): they are recomputed from the staged rest positions with the setup kernel's
//   own FP32 arithmetic (rest_shape), so they are bitwise the stored values.
// ---------------------------------------------------------------------------------------------
// One tile, all phases: stage the local nodes' positions and rest positions, evaluate the tets
// (phase 0 forces and energies; phase 1, when hd, the stiffness diagonal), gather each phase.
template <typename R, int MAT, bool F32M>
__device__ __forceinline__ void tile_body(const DevParams &p, int chk, const R *__restrict__ x, const TileDev &td,
                                          int b, const TileInfo &ti, int hdnow, const float *__restrict__ rest,
                                          R *__restrict__ fown, R *__restrict__ hpart, float *__restrict__ fownh,
                                          float *__restrict__ hparth, unsigned char *smem, double *lds_red) {
  const int L = ti.own_cnt + ti.halo_cnt, N = p.nn, tid = threadIdx.x;
  typedef typename std::conditional<F32M, float, R>::type C;
  R *sx = (R *)smem;
  C *rec = (C *)(sx + 4 * td.lcap);
  float4 *srest = (float4 *)(smem + tile_rest_offset<R>(td));
  const int nthr = blockDim.x;
  // neo-Hookean: tet tid's local ids and slots, loaded before the staging so that they arrive with
  // it (FP64 155 -> 148 us per launch at 6.3 M tets).  The stretch material's kernel would lose a
  // wave per SIMD to the 4 registers it holds across the staging (FP32 137 -> 147 us), so it loads
  // them after the barrier.
  constexpr bool PRE = MAT == 0;
  const uint4 lp0 = PRE && tid < ti.tet_cnt ? td.tet_lp[(size_t)ti.tet_start + tid] : make_uint4(0, 0, 0, 0);
  const C mu = (C)td.inst_mu[ti.inst], lam = (C)td.inst_lam[ti.inst];
  for (int i = tid; i < L; i += nthr) {
    const int g = i < ti.own_cnt ? ti.own_start + i : td.halo[ti.halo_off + i - ti.own_cnt];
    sx[4 * i] = x[g];
    sx[4 * i + 1] = x[N + g];
    sx[4 * i + 2] = x[2 * (size_t)N + g];
    srest[i] = make_float4(rest[g], rest[N + g], rest[2 * (size_t)N + g], 0.0f);
  }
  __syncthreads();
  double e_el = 0.0;
  // the tile's tets and their gather for one phase
  auto run_phase = [&](int phase) {
    for (int s = tid; s < ti.tet_cnt; s += nthr) {
      const uint4 lp = PRE && s == tid ? lp0 : td.tet_lp[(size_t)ti.tet_start + s];
      const int lid[4] = {(int)(lp.x & 0xffff), (int)(lp.x >> 16), (int)(lp.y & 0xffff), (int)(lp.y >> 16)};
      float r[4][3];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float4 q = srest[lid[c]];
        r[c][0] = q.x;
        r[c][1] = q.y;
        r[c][2] = q.z;
      }
      float Dmf[9], Vf;
      rest_shape(r, Dmf, Vf);
      tile_tet<R, C, MAT>(phase, lp, Dmf, Vf, sx, rec, mu, lam, chk, p.clamp, e_el);
    }
    __syncthreads();
    tile_gather<R>(phase, ti, td, rec, N, fown, hpart, fownh, hparth);
  };
  if constexpr (sizeof(C) == 8) {
    // FP64 arithmetic: the two phases as separate code.  As one loop over the phase the compiler
    // shares their common arithmetic and keeps its constants in registers, and the neo-Hookean
    // kernel needs 114 VGPRs (4 waves per SIMD); as separate code, 94 (5 waves; 144 -> 139 us
    // per launch at 6.3 M tets), and 78 with the logarithm's constants in scalar registers.
    run_phase(0);
    if (hdnow) {
      __syncthreads();  // the Hessian phase reuses the contribution slots
      run_phase(1);
    }
  } else {
    for (int phase = 0; phase <= hdnow; ++phase) {
      run_phase(phase);
      if (phase == 0 && hdnow) __syncthreads();
    }
  }
  if (chk) {
    const double se = block_sum(e_el, lds_red);
    if (tid == 0) td.tpart[b] = se;
  }
}

// F32M (FP64 state only): the per-tet arithmetic in FP32 after F (the mixed mode's early passes).
// The FP64 neo-Hookean kernel is held to 6 waves per SIMD (at most 80 VGPRs: it needs 78 without
// spilling), which pays with tiles small enough for 6 workgroups' LDS (224 tets, see pbs_create).
template <typename R, int MAT, bool F32M>
__global__ void __launch_bounds__(256, (sizeof(R) == 8 && MAT == 0 && !F32M) ? 6 : 1)
k_tile(DevParams p, const Ctrl *__restrict__ ctrl, Ring<R> ring, TileDev td, const float *__restrict__ rest,
       R *__restrict__ fown, R *__restrict__ hpart, float *__restrict__ fownh, float *__restrict__ hparth) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double lds_red[16];
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  // One instance: its control block is ctrl[0], read together with the tile descriptor rather than
  // after it, so the staging loads wait for one dependent memory access less.  Several: the
  // tile's instance's block, read once the descriptor has arrived.
  PassCtl c = pass_ctl(ctrl);
  const TileInfo ti = td.tiles[b];
  // both reads in flight before either is waited for (otherwise the compiler folds the first into
  // the second, which waits for the descriptor)
  asm volatile("" ::"s"(c.done), "s"(c.cur), "s"(c.hd_now), "s"(c.check_now), "s"(ti.inst));
  if (td.ninst != 1) c = pass_ctl(ctrl + ti.inst);
  if (c.done) return;
  tile_body<R, MAT, F32M>(p, c.check_now, ring.b[c.cur], td, b, ti, c.hd_now, rest, fown, hpart, fownh, hparth, smem,
                          lds_red);
}

// ---------------------------------------------------------------------------------------------
// Node pass (updateNodesNextIter + updateChebyshev, CUDAInterface.cu:572-623): force = owned
// partial + the node's halo partials in tile order; Jacobi-preconditioned step with Chebyshev.
// g = (x - x~) m/h^2 - f - s (m grav + f_ext), s = 1 (REF_GPU) or 2 (REF_CPU, DescentOptimize.cpp:51-52).
// ---------------------------------------------------------------------------------------------
// GRAD_ONLY: only the |g|^2 partials (the warm start's gradient passes); nothing else is written.
template <typename R, bool GRAD_ONLY>
__global__ void __launch_bounds__(256)
k_node(DevParams p, const Ctrl *__restrict__ ctrl, Ring<R> ring, TileDev td, const R *__restrict__ fown,
       const R *__restrict__ hpart, const float *__restrict__ fownh, const float *__restrict__ hparth,
       float *__restrict__ hd, const float *__restrict__ mnode, const float *__restrict__ ext,
       const R *__restrict__ xt) {
  __shared__ double lds_red[48];
  // One instance: control block ctrl[0], node blocks of 256 from node 0, read without waiting for
  // the block's instance.  Several: the block's instance first.
  PassCtl c = pass_ctl(ctrl);
  asm volatile("" ::"s"(c.done), "s"(c.cur), "s"(c.nxt), "s"(c.last), "s"(c.hd_now), "s"(c.cheb_now), "s"(c.alpha),
               "s"(c.omega_next));  // read now, not folded into the several-instance read below
  int n0 = 256 * (int)blockIdx.x, nend = p.nn;
  if (td.ninst != 1) {
    const int inst = td.blk_inst[blockIdx.x];
    c = pass_ctl(ctrl + inst);
    n0 = td.blk_start[blockIdx.x];
    nend = td.inst_node_off[inst + 1];
  }
  if (c.done) return;
  const int hdnow = c.hd_now, cheb = c.cheb_now;
  const R alpha = (R)c.alpha, om = (R)c.omega_next;
  const R *__restrict__ x = ring.b[c.cur];
  const R *__restrict__ xl = ring.b[c.last];
  R *__restrict__ xn = ring.b[c.nxt];
  const int n = n0 + (int)threadIdx.x;
  const int N = p.nn;
  double en = 0.0, g2 = 0.0, f2 = 0.0;
  if (n < nend) {
    // every load of the node that depends on no other is issued first (the flags are uniform), so
    // the thread waits for memory about twice (these loads, then its halo partials), not once per
    // array
    const int b0 = td.bnd_off[n], b1 = td.bnd_off[n + 1];
    const float m = mnode[n];  // the node's mass, or -1: fixed
    const bool fixed = m < 0.0f;
    R f[3], xc[3], xtv[3], xlv[3];
    float hv[3], ev[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const size_t i = (size_t)d * N + n;
      f[d] = fown[i];
      xc[d] = x[i];
      xtv[d] = xt[i];
      xlv[d] = (R)0;
      ev[d] = 0.0f;
      hv[d] = 0.0f;
    }
    if (!GRAD_ONLY && cheb)
#pragma unroll
      for (int d = 0; d < 3; ++d) xlv[d] = xl[(size_t)d * N + n];
    if (p.has_ext)
#pragma unroll
      for (int d = 0; d < 3; ++d) ev[d] = ext[(size_t)d * N + n];
    if (!GRAD_ONLY)
#pragma unroll
      for (int d = 0; d < 3; ++d) hv[d] = hdnow ? fownh[(size_t)d * N + n] : hd[(size_t)d * N + n];
    for (int e = b0; e < b1; ++e) {
      const R *hp = hpart + 3 * (size_t)e;
#pragma unroll
      for (int d = 0; d < 3; ++d) f[d] += hp[d];
    }
    if (!GRAD_ONLY && hdnow) {
      for (int e = b0; e < b1; ++e) {
        const float *hp = hparth + 3 * (size_t)e;
#pragma unroll
        for (int d = 0; d < 3; ++d) hv[d] += hp[d];
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) hd[(size_t)d * N + n] = hv[d];
    }
    if (fixed) {
      if (!GRAD_ONLY)
#pragma unroll
        for (int d = 0; d < 3; ++d) xn[(size_t)d * N + n] = xc[d];
    } else {
      const R hstep = (R)p.h;
      const R mh2 = (R)m / (hstep * hstep);
      R e = (R)0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const R dd = xc[d] - xtv[d];
        R fe = (R)(p.grav[d] * m);  // = the setup's gravity force (mass before the fixed override)
        if (p.has_ext) fe += (R)ev[d];
        const R gr = dd * mh2 - f[d] - (R)p.ext_scale * fe;
        e += (R)0.5 * dd * dd * mh2 - fe * xc[d];
        g2 += (double)gr * (double)gr;
        f2 += (double)f[d] * (double)f[d];
        if (!GRAD_ONLY) {
          R nx = xc[d] - alpha * gr / ((R)hv[d] + mh2);
          if (cheb) nx = om * (nx - xlv[d]) + xlv[d];
          xn[(size_t)d * N + n] = nx;
        }
      }
      if (!GRAD_ONLY) en += (double)e;
    }
  }
  if (!GRAD_ONLY && p.f32math) {  // mixed mode: also |f|^2, the scale of the FP32 force rounding
    double v[3] = {en, g2, f2};
    block_sums<3>(v, lds_red);
    if (threadIdx.x == 0) {
      td.npart[2 * blockIdx.x] = v[0];
      td.npart[2 * blockIdx.x + 1] = v[1];
      td.fpart[blockIdx.x] = v[2];
    }
  } else {
    double v[2] = {en, g2};
    block_sums<2>(v, lds_red);
    if (threadIdx.x == 0) {
      td.npart[2 * blockIdx.x] = v[0];
      td.npart[2 * blockIdx.x + 1] = v[1];
    }
  }
}

// acc[k] += v[K i + k] for i = first, first + 1024, ... below end, in that order: the loads of 8
// consecutive terms are issued before their additions (indices past the end are clamped to
// `first`, whose value is then not added), so the thread waits for memory once per 8 terms
// instead of once per term, and the sums are those of the plain loop.
template <int K>
__device__ __forceinline__ void strided_sum(const double *__restrict__ v, int first, int end, double *acc) {
  constexpr int U = 8;
  for (int base = first; base < end; base += U * 1024) {
    double t[U][K];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * 1024 < end ? base + u * 1024 : first;
#pragma unroll
      for (int k = 0; k < K; ++k) t[u][k] = v[(size_t)K * i + k];
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (base + u * 1024 < end)
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] += t[u][k];
  }
}

// Control step, one workgroup per instance: FP64 sums in a fixed order of the instance's tile
// energies (check passes) and node-block partials (energy, |g|^2), then the loop control of the
// reference.
__global__ void __launch_bounds__(1024) k_control(DevParams p, Ctrl *__restrict__ ctrl, TileDev td) {
  __shared__ double la[1024], lc[1024], lf[1024];
  const int inst = blockIdx.x;
  ctrl += inst;
  if (ctrl->done) return;
  const int tid = threadIdx.x;
  double a = 0.0, c = 0.0;
  if (ctrl->check_now) strided_sum<1>(td.tpart, td.inst_tile_off[inst] + tid, td.inst_tile_off[inst + 1], &a);
  {
    double ac[2] = {a, c};
    strided_sum<2>(td.npart, td.inst_blk_off[inst] + tid, td.inst_blk_off[inst + 1], ac);
    a = ac[0];
    c = ac[1];
  }
  const bool mix = p.f32math != 0;  // mixed mode: also |f|^2 of this pass, the FP32 force scale
  double f = 0.0;
  if (mix) strided_sum<1>(td.fpart, td.inst_blk_off[inst] + tid, td.inst_blk_off[inst + 1], &f);
  la[tid] = a;
  lc[tid] = c;
  lf[tid] = f;
  __syncthreads();
  // pairwise tree: levels 512 .. 64 through LDS, levels 32 .. 1 inside wavefront 0 by shuffles,
  // which pair lane t with lane t + st exactly as the LDS levels pair slots
  for (int st = 512; st >= 64; st >>= 1) {
    if (tid < st) {
      la[tid] += la[tid + st];
      lc[tid] += lc[tid + st];
      if (mix) lf[tid] += lf[tid + st];
    }
    __syncthreads();
  }
  if (tid < 64) {
    double E = la[tid], g2 = lc[tid], f2 = lf[tid];
#pragma unroll
    for (int st = 32; st > 0; st >>= 1) {
      E += __shfl_down(E, st, 64);
      g2 += __shfl_down(g2, st, 64);
      if (mix) f2 += __shfl_down(f2, st, 64);
    }
    if (tid == 0) {
      Ctrl C = *ctrl;
      if (p.f32math) C.fnorm = sqrt(f2);
      control_apply(p, C, E, g2);
      *ctrl = C;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Step start (initDescentOpt, CUDAInterface.cu:482-493; DescentOptimize.cpp:35-41) and the
// device loop state reset.
// ---------------------------------------------------------------------------------------------
template <typename R>
__global__ void k_step_init(DevParams p, Ctrl *__restrict__ ctrl, int ninst, Ring<R> ring, const R *__restrict__ xpos,
                            const R *__restrict__ v, const R *__restrict__ vl, R *__restrict__ xt) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < p.nn) {
    const R h = (R)p.h;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const size_t i = (size_t)d * p.nn + n;
      const R t = xpos[i] + v[i] * h;
      xt[i] = t;
      R x0;
      if (p.semantics == 1) x0 = t + (R)(p.init_momentum * p.h) * (v[i] - vl[i]);
      else if (sizeof(R) == 4) x0 = (R)((double)t + (double)p.init_momentum * (double)(v[i] - vl[i]) * (double)p.h);
      else x0 = t + (R)p.init_momentum * (v[i] - vl[i]) * h;
      ring.b[0][i] = x0;
    }
  }
  if (n < ninst) {  // one control block per instance
    Ctrl C;
    C.k = 0;
    C.budget = p.iterations;
    C.passes = 0;
    C.rollbacks = 0;
    C.converged = 0;
    C.used_learned = 0;
    C.first = 1;
    C.cur = 0;
    C.nxt = 1;
    C.last = 2;
    C.old = 3;
    C.hd_now = 1;
    C.check_now = 1;
    C.kc = 0;
    C.alpha = p.alpha0;
    C.rho = 0.0f;
    C.omega = 1.0f;
    float rho = 0.0f, om = 1.0f;
    cheb_schedule(p, 0, rho, om);
    C.rho_next = rho;
    C.omega_next = om;
    C.cheb_now = om != 1.0f;
    C.ws_beta = 0.0f;
    C.E = 0.0;
    C.E_old = 0.0;
    C.g0 = -1.0;
    C.g = 0.0;
    C.g_plain = 0.0;
    C.g_learned = 0.0;
    C.done = (p.iterations <= 0 || p.alpha0 < p.alpha_min) ? 1 : 0;
    ctrl[n] = C;
  }
}

// Step end: velocities (updateVelocity, CUDAInterface.cu:625-639, or DescentOptimize.cpp:183-185)
// and the new positions.
template <typename R>
__global__ void k_step_final(DevParams p, const Ctrl *__restrict__ ctrl, TileDev td, Ring<R> ring, R *__restrict__ xpos,
                             R *__restrict__ v, R *__restrict__ vl, const R *__restrict__ xt,
                             const int *__restrict__ mask) {
  const int inst = td.blk_inst[blockIdx.x];
  const int n = td.blk_start[blockIdx.x] + (int)threadIdx.x;
  if (n >= td.inst_node_off[inst + 1]) return;
  const R *__restrict__ x = ring.b[ctrl[inst].cur];
  const R h = (R)p.h;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const size_t i = (size_t)d * p.nn + n;
    const R xf = x[i];
    if (p.semantics == 1) {
      vl[i] = v[i];
      v[i] = (R)p.damping * (xf - xpos[i]) / h;
    } else if (!mask[n]) {
      vl[i] = v[i];
      const R vv = v[i] + (xf - xt[i]) * ((R)1 / h);
      v[i] = (sizeof(R) == 4) ? (R)((double)vv * (double)p.damping) : vv * (R)p.damping;
    }
    xpos[i] = xf;
  }
}

// Control block for a standalone force / Hessian evaluation on ring buffer `cur` (hd, chk: the
// tile pass's Hessian phase and energy partials).
__global__ void k_standalone_ctrl(Ctrl *ctrl, int cur, int hd, int chk) {
  Ctrl C = {};
  C.cur = cur;
  C.nxt = (cur + 1) & 3;
  C.last = (cur + 2) & 3;
  C.old = (cur + 3) & 3;
  C.hd_now = hd;
  C.check_now = chk;
  C.done = 0;
  *ctrl = C;
}

// Per-node sums of the tile pass's partials, in the node pass's order (owned partial, then the
// halo partials in tile order), written as doubles in the caller's node order (AoS): the elastic
// forces (which = 0) or the Hessian diagonal (which = 1) the descent uses.
template <typename R>
__global__ void k_sum_partials(TileDev td, int N, int which, const R *__restrict__ fown, const R *__restrict__ hpart,
                               const float *__restrict__ fownh, const float *__restrict__ hparth,
                               const int *__restrict__ perm, double *__restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int b0 = td.bnd_off[n], b1 = td.bnd_off[n + 1];
  const size_t o = 3 * (size_t)perm[n];
  if (which == 0) {
    R f[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) f[d] = fown[(size_t)d * N + n];
    for (int e = b0; e < b1; ++e)
#pragma unroll
      for (int d = 0; d < 3; ++d) f[d] += hpart[3 * (size_t)e + d];
#pragma unroll
    for (int d = 0; d < 3; ++d) out[o + d] = (double)f[d];
  } else {
    float h[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) h[d] = fownh[(size_t)d * N + n];
    for (int e = b0; e < b1; ++e)
#pragma unroll
      for (int d = 0; d < 3; ++d) h[d] += hparth[3 * (size_t)e + d];
#pragma unroll
    for (int d = 0; d < 3; ++d) out[o + d] = (double)h[d];
  }
}

// Gather of per-tet records into per-node planes (elastic forces or Hessian diagonal).
template <typename T>
__global__ void k_gather(int N, const int *__restrict__ inc_off, const int *__restrict__ inc,
                         const T *__restrict__ rec, T *__restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  T a = 0, b = 0, c = 0;
  for (int e = inc_off[n]; e < inc_off[n + 1]; ++e) {
    const T *r = rec + 3 * (size_t)inc[e];
    a += r[0];
    b += r[1];
    c += r[2];
  }
  out[n] = a;
  out[N + n] = b;
  out[2 * (size_t)N + n] = c;
}

// AoS (caller's node order) <-> planes (solver's node order); perm maps solver -> caller index.
template <typename A, typename B>
__global__ void k_aos_to_planes(const A *__restrict__ aos, B *__restrict__ planes, const int *__restrict__ perm, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t o = 3 * (size_t)(perm ? perm[i] : i);
#pragma unroll
  for (int d = 0; d < 3; ++d) planes[(size_t)d * n + i] = (B)aos[o + d];
}

template <typename A, typename B>
__global__ void k_planes_to_aos(const A *__restrict__ planes, B *__restrict__ aos, const int *__restrict__ perm, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t o = 3 * (size_t)(perm ? perm[i] : i);
#pragma unroll
  for (int d = 0; d < 3; ++d) aos[o + d] = (B)planes[(size_t)d * n + i];
}

__global__ void k_fill(float *p, float v, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------------------------
// Launch wrappers
// ---------------------------------------------------------------------------------------------
static inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

// A launch that, given events, records its own start and end in them.
template <typename F, typename... Args>
static void launch_timed(F kernel, dim3 grid, dim3 block, size_t lds, hipStream_t st, hipEvent_t e0, hipEvent_t e1,
                         Args... args) {
  if (e0) hipExtLaunchKernelGGL(kernel, grid, block, (std::uint32_t)lds, st, e0, e1, 0u, args...);
  else hipLaunchKernelGGL(kernel, grid, block, lds, st, args...);
}

template <typename R>
void Launch<R>::setup_tets(const DevMesh<R> &m, const DevParams &p, int bs, hipStream_t st) {
  if (p.nt > 0) hipLaunchKernelGGL(k_setup_tets, dim3(nblk(p.nt, bs)), dim3(bs), 0, st, p, m.rest, m.tets, m.dmi, m.vol, m.flag);
}
template <typename R>
void Launch<R>::setup_nodes(const DevMesh<R> &m, const DevParams &p, int bs, hipStream_t st) {
  hipLaunchKernelGGL(k_setup_nodes, dim3(nblk(p.nn, bs)), dim3(bs), 0, st, p, m.inc_off, m.inc, m.vol, m.mask, m.mass,
                     m.mnode, m.gforce);
}
template <typename R>
void Launch<R>::tet(const DevMesh<R> &m, const TileDev &td, const DevParams &p, int bs, int force_hd, int write_e,
                    hipStream_t st) {
  if (p.nt <= 0) return;
  R *te = write_e ? m.tete : nullptr;
  if (p.material == 0)
    hipLaunchKernelGGL((k_tet<R, 0>), dim3(nblk(p.nt, bs)), dim3(bs), 0, st, p, m.ctrl, m.ring, m.tets, m.dmi, m.vol, m.frec, m.hrec, m.etet, te, force_hd, td);
  else
    hipLaunchKernelGGL((k_tet<R, 1>), dim3(nblk(p.nt, bs)), dim3(bs), 0, st, p, m.ctrl, m.ring, m.tets, m.dmi, m.vol, m.frec, m.hrec, m.etet, te, force_hd, td);
}
template <typename R>
size_t Launch<R>::tile_lds(const TileDev &td) {
  return tile_rest_offset<R>(td) + 16 * (size_t)td.lcap;
}
template <typename R>
void Launch<R>::tile_pass(const DevMesh<R> &m, const TileDev &td, const DevParams &p, hipStream_t st, hipEvent_t e0,
                          hipEvent_t e1) {
  const size_t lds = tile_lds(td);
  const bool f32m = sizeof(R) == 8 && p.f32math;
#define PBS_TILE_LAUNCH(MAT, F32M)                                                                               \
  launch_timed(k_tile<R, MAT, F32M>, dim3(td.ntiles), dim3(td.threads), lds, st, e0, e1, p, m.ctrl, m.ring, td, \
               m.rest, m.fown, m.hpart, m.fownh, m.hparth)
  if (p.material == 0 && !f32m) PBS_TILE_LAUNCH(0, false);
  else if (p.material == 0) PBS_TILE_LAUNCH(0, true);
  else if (!f32m) PBS_TILE_LAUNCH(1, false);
  else PBS_TILE_LAUNCH(1, true);
#undef PBS_TILE_LAUNCH
}
template <typename R>
void Launch<R>::node_pass(const DevMesh<R> &m, const TileDev &td, const DevParams &p, hipStream_t st, hipEvent_t e0,
                          hipEvent_t e1) {
  launch_timed(k_node<R, false>, dim3(td.nblk), dim3(256), 0, st, e0, e1, p, m.ctrl, m.ring, td, m.fown, m.hpart,
               m.fownh, m.hparth, m.hd, m.mnode, m.ext, m.xt);
}
template <typename R>
void Launch<R>::grad_pass(const DevMesh<R> &m, const TileDev &td, const DevParams &p, hipStream_t st) {
  hipLaunchKernelGGL((k_node<R, true>), dim3(td.nblk), dim3(256), 0, st, p, m.ctrl, m.ring, td, m.fown, m.hpart,
                     m.fownh, m.hparth, m.hd, m.mnode, m.ext, m.xt);
}
template <typename R>
void Launch<R>::control(const DevMesh<R> &m, const TileDev &td, const DevParams &p, hipStream_t st, hipEvent_t e0,
                        hipEvent_t e1) {
  launch_timed(k_control, dim3(td.ninst), dim3(1024), 0, st, e0, e1, p, m.ctrl, td);
}
template <typename R>
void Launch<R>::step_init(const DevMesh<R> &m, const TileDev &td, const DevParams &p, int bs, hipStream_t st) {
  hipLaunchKernelGGL((k_step_init<R>), dim3(nblk(std::max(p.nn, td.ninst), bs)), dim3(bs), 0, st, p, m.ctrl, td.ninst,
                     m.ring, m.xpos, m.v, m.vl, m.xt);
}
template <typename R>
void Launch<R>::step_final(const DevMesh<R> &m, const TileDev &td, const DevParams &p, hipStream_t st) {
  hipLaunchKernelGGL((k_step_final<R>), dim3(td.nblk), dim3(256), 0, st, p, m.ctrl, td, m.ring, m.xpos, m.v, m.vl, m.xt,
                     m.mask);
}
template <typename R>
void Launch<R>::standalone_ctrl(const DevMesh<R> &m, int cur, hipStream_t st, int hd, int chk) {
  hipLaunchKernelGGL(k_standalone_ctrl, dim3(1), dim3(1), 0, st, m.ctrl, cur, hd, chk);
}
template <typename R>
void Launch<R>::sum_partials(const DevMesh<R> &m, const TileDev &td, const DevParams &p, int which, double *out,
                             hipStream_t st) {
  hipLaunchKernelGGL((k_sum_partials<R>), dim3(nblk(p.nn, 256)), dim3(256), 0, st, td, p.nn, which, m.fown, m.hpart,
                     m.fownh, m.hparth, m.perm, out);
}
template <typename R>
void Launch<R>::gather(const DevMesh<R> &m, const DevParams &p, int bs, int which, void *out, hipStream_t st) {
  if (which == 0)
    hipLaunchKernelGGL((k_gather<R>), dim3(nblk(p.nn, bs)), dim3(bs), 0, st, p.nn, m.inc_off, m.inc, m.frec, (R *)out);
  else
    hipLaunchKernelGGL((k_gather<float>), dim3(nblk(p.nn, bs)), dim3(bs), 0, st, p.nn, m.inc_off, m.inc, m.hrec, (float *)out);
}
template <typename R>
void Launch<R>::aos_to_planes(const float *aos, R *planes, const int *perm, int n, hipStream_t st) {
  hipLaunchKernelGGL((k_aos_to_planes<float, R>), dim3(nblk(n, 256)), dim3(256), 0, st, aos, planes, perm, n);
}
template <typename R>
void Launch<R>::planes_to_aos(const R *planes, float *aos, const int *perm, int n, hipStream_t st) {
  hipLaunchKernelGGL((k_planes_to_aos<R, float>), dim3(nblk(n, 256)), dim3(256), 0, st, planes, aos, perm, n);
}
template <typename R>
void Launch<R>::planes_to_aos64(const R *planes, double *aos, const int *perm, int n, hipStream_t st) {
  hipLaunchKernelGGL((k_planes_to_aos<R, double>), dim3(nblk(n, 256)), dim3(256), 0, st, planes, aos, perm, n);
}
template <typename R>
void Launch<R>::aos64_to_planes(const double *aos, R *planes, const int *perm, int n, hipStream_t st) {
  hipLaunchKernelGGL((k_aos_to_planes<double, R>), dim3(nblk(n, 256)), dim3(256), 0, st, aos, planes, perm, n);
}
void aos_to_planes_f(const float *aos, float *planes, const int *perm, int n, hipStream_t st) {
  hipLaunchKernelGGL((k_aos_to_planes<float, float>), dim3(nblk(n, 256)), dim3(256), 0, st, aos, planes, perm, n);
}
void planes_to_aos_f(const float *planes, float *aos, const int *perm, int n, hipStream_t st) {
  hipLaunchKernelGGL((k_planes_to_aos<float, float>), dim3(nblk(n, 256)), dim3(256), 0, st, planes, aos, perm, n);
}

void fill_f(float *p, float v, long n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_fill, dim3(nblk(n, 256)), dim3(256), 0, st, p, v, n);
}

template struct Launch<float>;
template struct Launch<double>;

}  // namespace pbs
