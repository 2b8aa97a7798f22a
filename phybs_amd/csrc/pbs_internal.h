// pbs_internal.h -- structures shared by the kernels (pbs_kernels.hip) and the host driver
// (pbs_capi.cpp).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbs_tiles.h"

namespace pbs {

// Kernel-side copy of the solver constants (passed by value to every kernel).
struct DevParams {
  int nn, nt;
  float h, mu, lam, density, fixed_mass;
  float grav[3];
  float alpha0, alpha_decay, alpha_min;
  int hd_refresh, energy_check;
  float damping, init_momentum;
  int profile_n;
  int profile_k[8];
  float profile_rho[8];
  int semantics;  // 0 REF_GPU, 1 REF_CPU
  int material;   // 0 neo-Hookean, 1 stretch
  int clamp;      // 1 max(E,0), 2 E>1e-6?E:0
  float ext_scale;
  int has_ext;
  int iterations;
  double relgrad_tol;
  int f32math;  // FP64 state: the tile pass evaluates the per-tet terms in FP32 (mixed mode)
};

// Device-resident loop state.  The reference keeps k, alpha, the energies and the Chebyshev
// scalars on the host and copies node arrays to roll back (CUDAInterface.cu:911-1003); here
// the control kernel decides on the device and rollbacks/swaps are buffer-role changes.
struct Ctrl {
  int k, budget, done, passes;
  int rollbacks, converged, used_learned, first;
  int cur, nxt, last, old;  // roles of the 4 position buffers
  int hd_now, check_now, cheb_now, kc;  // kc: iteration of the last Chebyshev restart
  float alpha, rho, omega, rho_next, omega_next, ws_beta;  // ws_beta: warm start's chosen step
  double E, E_old, g0, g;
  double g_plain, g_learned;  // warm start: gradient norms of the two initial guesses
  double fnorm;               // mixed mode: |f| (internal forces on free nodes) of the last FP32 pass
};

// Tile partition on the device (see pbs_tiles.h for the host-side construction).
struct TileDev {
  int ntiles, lcap, tcap, threads;  // threads: workgroup size of the tile pass (64..256)
  // instances (independent meshes; 1 unless batched): node / tile / tet ranges, node blocks of
  // 256 nodes per instance (block b: instance blk_inst[b], first node blk_start[b]), Lame constants
  int ninst, nblk;
  const int *inst_node_off, *inst_tile_off, *inst_tet_off, *inst_blk_off;
  const int *blk_inst, *blk_start;
  const float *inst_mu, *inst_lam;
  long halo_total;
  const TileInfo *tiles;
  const int *halo;
  const uint4 *tet_lp;          // per tet: corner local ids, corner contribution positions (16 bits each)
  const unsigned short *tile_off;
  const int *bnd_off;           // per node: its halo partial slots [bnd_off[n], bnd_off[n+1]), tile order
  const int *halo_slot;         // per halo entry: its partial slot
  double *tpart;                // per tile: elastic energy of its tets
  double *npart;                // per node block: node energy, |g|^2
  double *fpart;                // per node block, mixed-mode FP32 passes only: |f|^2
};

// Values per corner contribution in the tile pass's LDS: 4 floats (one 16-byte access) or 3 doubles.
template <typename R> struct ContribWidth { static constexpr int value = 3; };
template <> struct ContribWidth<float> { static constexpr int value = 4; };

template <typename R> struct Ring {
  R *b[4];  // each: 3 planes of nn values (x | y | z)
};

// Device buffers of one solver instance.
template <typename R> struct DevMesh {
  int4 *tets;       // T (solver order, solver node ids)
  float *rest;      // 3 planes
  float *dmi;       // 9 planes of T (Dm^-1 row-major entry e at dmi[e*T + t])
  float *vol;       // T
  int *inc_off;     // N+1
  int *inc;         // 4T, entry = 4*tet + local, per node in ascending caller tet id
  int *perm;        // N: solver node -> caller node
  int *mask;        // N
  float *mass;      // N (fixed nodes: fixed_mass)
  float *mnode;     // N: the computed mass of a free node, -1 for a fixed node (the node pass reads
                    // this one array instead of mask and mass)
  float *gforce;    // 3 planes: mass (before the fixed override) * grav
  float *ext;       // 3 planes
  float *hd;        // 3 planes: Hessian diagonal
  R *frec;          // 12T per-tet force records
  float *hrec;      // 12T per-tet Hessian-diagonal records
  R *xt;            // 3 planes: inertial target x~
  R *xpos;          // 3 planes: positions at step start / after the step
  R *v, *vl;        // 3 planes each
  Ring<R> ring;
  double *etet;     // per tet-block energy partials
  double *enode;    // per node-block energy partials
  double *gpart;    // per node-block |g|^2 partials
  Ctrl *ctrl;
  int *flag;        // setup error flag
  R *tete;          // T, optional per-tet energy output
  R *fown, *hpart;  // tile pass outputs: owned-node partial forces (3 planes of N), halo partials (3 x H)
  float *fownh, *hparth;  // the same for the Hessian diagonal
};

// ---- launch wrappers (pbs_kernels.hip) ----
template <typename R> struct Launch {
  static void setup_tets(const DevMesh<R> &m, const DevParams &p, int bs, hipStream_t st);
  static void setup_nodes(const DevMesh<R> &m, const DevParams &p, int bs, hipStream_t st);
  static void tet(const DevMesh<R> &m, const TileDev &td, const DevParams &p, int bs, int force_hd, int write_e,
                  hipStream_t st);
  static size_t tile_lds(const TileDev &td);
  // The kernels of a descent pass.  With events e0, e1 the launch records its own start and end
  // in them (hipExtLaunchKernelGGL: the dispatch's timestamps, no extra stream operations).
  static void tile_pass(const DevMesh<R> &m, const TileDev &td, const DevParams &p, hipStream_t st,
                        hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
  static void node_pass(const DevMesh<R> &m, const TileDev &td, const DevParams &p, hipStream_t st,
                        hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
  static void grad_pass(const DevMesh<R> &m, const TileDev &td, const DevParams &p, hipStream_t st);
  static void control(const DevMesh<R> &m, const TileDev &td, const DevParams &p, hipStream_t st,
                      hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
  static void step_init(const DevMesh<R> &m, const TileDev &td, const DevParams &p, int bs, hipStream_t st);
  static void step_final(const DevMesh<R> &m, const TileDev &td, const DevParams &p, hipStream_t st);
  static void standalone_ctrl(const DevMesh<R> &m, int cur, hipStream_t st, int hd = 0, int chk = 0);
  // per-node sums of the tile pass's partials (which: 0 forces, 1 Hessian diagonal), doubles in
  // the caller's node order (3N AoS)
  static void sum_partials(const DevMesh<R> &m, const TileDev &td, const DevParams &p, int which, double *out,
                           hipStream_t st);
  static void gather(const DevMesh<R> &m, const DevParams &p, int bs, int which, void *out_planes, hipStream_t st);
  static void aos_to_planes(const float *aos, R *planes, const int *perm, int n, hipStream_t st);
  static void planes_to_aos(const R *planes, float *aos, const int *perm, int n, hipStream_t st);
  static void planes_to_aos64(const R *planes, double *aos, const int *perm, int n, hipStream_t st);
  static void aos64_to_planes(const double *aos, R *planes, const int *perm, int n, hipStream_t st);
};
// ---- learned warm start (pbs_learn.hip) ----
// Per-vertex network 18 -> 64 -> 64 -> 64 -> 3, SiLU; packed FP32 parameters W1 (18x64, in x out),
// b1, W2 (64x64), b2, W3, b3, W4 (64x3), b4.
constexpr int MLP_NPARAM = 18 * 64 + 64 + 2 * (64 * 64 + 64) + 64 * 3 + 3;  // 9731
constexpr int MLP_NSCALE = 7;       // 6 feature-group scales, 1 output scale
constexpr int MLP_WA = 11264;       // BF16 weights arranged as MFMA A operands (all 4 layers)
struct MlpDev {
  const float *params;         // packed FP32 master parameters
  const unsigned short *wA;    // BF16 A-operand arrangement of the weights
  float inv_s[MLP_NSCALE];     // float32(1) / scale
};
template <typename R> struct LearnLaunch {
  // features of every node at the current step start (x = xpos, v) into feat (N x 18, FP32), and
  // the inertial target y = x + h v + h^2 g (planes) into y
  static void features(const DevMesh<R> &m, const DevParams &p, const int *adj_off, const int *adj, float *feat,
                       R *y, hipStream_t st);
  // learned guess y + out * scale into ring.b[1] for free nodes (fixed nodes copy ring.b[0])
  static void guess(const DevMesh<R> &m, const DevParams &p, const MlpDev &net, const float *feat, const R *y,
                    hipStream_t st);
  // correction d = ring.b[1] - ring.b[0], smoothed by iters sweeps in FP32 (sm: 2 x N float4
  // scratch), then the selection: the gradient norm of the plain guess ring.b[0] and of
  // ring.b[0] + beta d for each beta; the descent starts from the best of them
  static void select(const DevMesh<R> &m, const TileDev &td, const DevParams &p, const int *adj_off, const int *adj,
                     const int *sell_off, const int *sell_col, int iters, const float *betas, int nbeta, float *sm,
                     hipStream_t st);
  // appends (features, x_new - y) of the free nodes of rank r with r % stride == phase to the
  // sample buffers (rank r at base + r / stride)
  static void collect(const DevMesh<R> &m, const DevParams &p, const float *feat, const R *y, const int *free_rank,
                      long base, int stride, int phase, float *ds_x, float *ds_y, hipStream_t st);
};
void mlp_arrange(const float *params, unsigned short *wA, hipStream_t st);
void mlp_forward(const MlpDev &net, const float *feat, int n, float *out, hipStream_t st);
// training: per 32-sample workgroup partial gradients (MLP_NPARAM each) and loss sums
void mlp_train_partials(const MlpDev &net, const float *ds_x, const float *ds_y, const int *idx, int n, int batch,
                        float *grad_part, float *loss_part, hipStream_t st);
void mlp_reduce(const float *grad_part, const float *loss_part, int nwg, float *grad, double *loss_acc, hipStream_t st);
void mlp_adam(float *params, float *m1, float *m2, const float *grad, float lr, float b1t, float b2t, hipStream_t st);

void aos_to_planes_f(const float *aos, float *planes, const int *perm, int n, hipStream_t st);
void planes_to_aos_f(const float *planes, float *aos, const int *perm, int n, hipStream_t st);
void fill_f(float *p, float v, long n, hipStream_t st);

}  // namespace pbs
