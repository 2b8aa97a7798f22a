// pbs_learn.hip -- kernels of the learned warm start (no reference counterpart; PhyBS has no
// learned component).
//
// A per-vertex network (18 features -> 64 -> 64 -> 64 -> 3, SiLU) predicts the displacement of
// the converged step from the inertial target y = x + h v + h^2 g.  Inference runs on the matrix
// cores: BF16 weights and activations, FP32 accumulation, v_mfma_f32_16x16x16_bf16 with the
// layers written transposed (H_{l+1}^T = act(W_l^T H_l^T + b)), so that the accumulator tile of one
// layer is already the B-operand tile of the next and nothing is shuffled through LDS.  Training
// (forward and backward in FP32 on the vector units, with the same BF16 roundings) produces
// per-workgroup gradient partials that are reduced in a fixed order; Adam updates FP32 master
// weights.
//
// The warm start is a guess, never a replacement: both the plain guess and the learned one are
// evaluated with one gradient pass each, and the descent starts from the learned guess only if its
// gradient norm is not larger (the plain norm stays the rel-grad denominator).
#include "pbs_device.h"
#include "pbs_internal.h"

namespace pbs {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest, ties to even.
__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ float bfround(float f) { return bf2f(f2bf(f)); }
__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// Layer table: inputs, outputs, 16-row output tiles, 16-wide input blocks, offsets of the
// A-operand arrangement, of the packed weights and of the packed biases.
struct Layer {
  int in, out, mt, kb, wa, pw, pb;
};
__constant__ Layer kLayers[4] = {{18, 64, 4, 2, 0, 0, 1152},
                                 {64, 64, 4, 4, 2048, 1216, 5312},
                                 {64, 64, 4, 4, 6144, 5376, 9472},
                                 {64, 3, 1, 4, 10240, 9536, 9728}};

// A operand of tile (mt, kb) for lane (q, r) = 4 BF16: W^T[16 mt + r][16 kb + 4 q + i], i = 0..3,
// i.e. W[16 kb + 4 q + i][16 mt + r], zero outside the layer.
__global__ void k_mlp_arrange(const float *__restrict__ params, unsigned short *__restrict__ wA) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= MLP_WA) return;
  int l = 3;
  while (l > 0 && e < kLayers[l].wa) --l;
  const Layer L = kLayers[l];
  const int el = e - L.wa, blk = el >> 8, lane = (el >> 2) & 63, i = el & 3;
  const int mt = blk / L.kb, kb = blk % L.kb, q = lane >> 4, r = lane & 15;
  const int o = 16 * mt + r, k = 16 * kb + 4 * q + i;
  wA[e] = (o < L.out && k < L.in) ? f2bf(params[L.pw + k * L.out + o]) : (unsigned short)0;
}

// Forward pass of one 16-sample tile by one wave.  feat: FP32 features (18 per sample); samples
// base .. base+15 (those >= n read as zero).  On return, lanes 0..15 hold the three normalised
// outputs of sample base + lane in o[0..2].
__device__ __forceinline__ void mlp_tile(const unsigned short *__restrict__ sW, const MlpDev &net,
                                         const float *__restrict__ feat, int base, int n, float o[3]) {
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
  const int sidx = base + r;
  const bool valid = sidx < n;
  bf16x4 x[2];
#pragma unroll
  for (int kb = 0; kb < 2; ++kb)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = 16 * kb + 4 * q + i;
      const float f = (valid && k < 18) ? feat[(size_t)sidx * 18 + k] * net.inv_s[k / 3] : 0.0f;
      x[kb][i] = (short)f2bf(f);
    }
  bf16x4 h[4], hn[4];
  const float *P = net.params;
  // layer 1: 18 (2 blocks) -> 64
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    f32x4 c;
#pragma unroll
    for (int i = 0; i < 4; ++i) c[i] = P[1152 + 16 * mt + 4 * q + i];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
      const bf16x4 a = *(const bf16x4 *)(sW + 0 + (mt * 2 + kb) * 256 + lane * 4);
      c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, x[kb], c, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) h[mt][i] = (short)f2bf(silu(c[i]));
  }
  // layers 2 and 3: 64 -> 64
#pragma unroll
  for (int l = 1; l < 3; ++l) {
    const int wa = l == 1 ? 2048 : 6144, pb = l == 1 ? 5312 : 9472;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      f32x4 c;
#pragma unroll
      for (int i = 0; i < 4; ++i) c[i] = P[pb + 16 * mt + 4 * q + i];
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) {
        const bf16x4 a = *(const bf16x4 *)(sW + wa + (mt * 4 + kb) * 256 + lane * 4);
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, h[kb], c, 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) hn[mt][i] = (short)f2bf(silu(c[i]));
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) h[mt] = hn[mt];
  }
  // layer 4: 64 -> 3 (rows 3..15 of the tile are zero weights)
  f32x4 c;
#pragma unroll
  for (int i = 0; i < 4; ++i) c[i] = (4 * q + i < 3) ? P[9728 + 4 * q + i] : 0.0f;
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    const bf16x4 a = *(const bf16x4 *)(sW + 10240 + kb * 256 + lane * 4);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, h[kb], c, 0, 0, 0);
  }
  o[0] = c[0];
  o[1] = c[1];
  o[2] = c[2];
}

__device__ __forceinline__ void load_wA(const unsigned short *__restrict__ wA, unsigned short *sW) {
  const uint4 *src = (const uint4 *)wA;
  uint4 *dst = (uint4 *)sW;
  for (int i = threadIdx.x; i < MLP_WA / 8; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
}

__global__ void __launch_bounds__(256) k_mlp_forward(MlpDev net, const float *__restrict__ feat, int n,
                                                     float *__restrict__ out) {
  __shared__ __align__(16) unsigned short sW[MLP_WA];
  load_wA(net.wA, sW);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int tile = blockIdx.x * 4 + wave; tile * 16 < n; tile += gridDim.x * 4) {
    float o[3];
    mlp_tile(sW, net, feat, tile * 16, n, o);
    const int sidx = tile * 16 + lane;
    if (lane < 16 && sidx < n) {
      out[(size_t)sidx * 3 + 0] = o[0];
      out[(size_t)sidx * 3 + 1] = o[1];
      out[(size_t)sidx * 3 + 2] = o[2];
    }
  }
}

// Features at the step start (x = current positions, v = current velocities):
//   u = x - X, h v, s = h v + h^2 g, and the 1-ring means of (u, h v, s) minus the node's own;
// fixed nodes get zeros.  Also writes the inertial target y = x + s.
template <typename R>
__global__ void k_features(DevParams p, const R *__restrict__ x, const R *__restrict__ v, const float *__restrict__ rest,
                           const int *__restrict__ mask, const int *__restrict__ adj_off, const int *__restrict__ adj,
                           float *__restrict__ feat, R *__restrict__ y) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = p.nn;
  if (n >= N) return;
  const R h = (R)p.h;
  R u[3], hv[3], s[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const size_t i = (size_t)d * N + n;
    u[d] = x[i] - (R)rest[i];
    hv[d] = h * v[i];
    s[d] = hv[d] + h * h * (R)p.grav[d];
    y[i] = x[i] + s[d];
  }
  float *f = feat + (size_t)n * 18;
  if (mask[n]) {
#pragma unroll
    for (int k = 0; k < 18; ++k) f[k] = 0.0f;
    return;
  }
  R mu[3] = {0, 0, 0}, mv[3] = {0, 0, 0}, ms[3] = {0, 0, 0};
  const int e0 = adj_off[n], e1 = adj_off[n + 1];
  for (int e = e0; e < e1; ++e) {
    const int m = adj[e];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const size_t i = (size_t)d * N + m;
      const R hvm = h * v[i];
      mu[d] += x[i] - (R)rest[i];
      mv[d] += hvm;
      ms[d] += hvm + h * h * (R)p.grav[d];
    }
  }
  const R c = (R)(e1 - e0 > 0 ? e1 - e0 : 1);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    f[d] = (float)u[d];
    f[3 + d] = (float)hv[d];
    f[6 + d] = (float)s[d];
    f[9 + d] = (float)(mu[d] / c - u[d]);
    f[12 + d] = (float)(mv[d] / c - hv[d]);
    f[15 + d] = (float)(ms[d] / c - s[d]);
  }
}

// Learned guess: ring.b[1] = y + out * scale for free nodes, ring.b[0] for fixed nodes.
template <typename R>
__global__ void __launch_bounds__(256) k_guess(DevParams p, MlpDev net, const float *__restrict__ feat,
                                               const R *__restrict__ y, const int *__restrict__ mask,
                                               const R *__restrict__ plain, R *__restrict__ out, float s_out) {
  __shared__ __align__(16) unsigned short sW[MLP_WA];
  load_wA(net.wA, sW);
  const int N = p.nn;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int tile = blockIdx.x * 4 + wave; tile * 16 < N; tile += gridDim.x * 4) {
    float o[3];
    mlp_tile(sW, net, feat, tile * 16, N, o);
    const int n = tile * 16 + lane;
    if (lane < 16 && n < N) {
      const bool fixed = mask[n] != 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const size_t i = (size_t)d * N + n;
        out[i] = fixed ? plain[i] : y[i] + (R)(o[d] * s_out);
      }
    }
  }
}

// Smoothing of the learned correction d = learned - plain over the node graph (fixed nodes pinned
// at 0): d'[n] = d[n] / 2 + (mean of d over the 1-ring) / 2.  The network's per-node errors are
// uncorrelated between neighbours, and such rough errors dominate the gradient norm through the
// stiff elastic forces; a few averaging sweeps keep the smooth part of the correction.  The
// correction is a heuristic guess (the gradient test decides whether it is used), so the sweeps
// run on FP32 xyz(w) quadruples: one 16-byte load per neighbour instead of three plane loads.
template <typename R>
__global__ void k_corr_diff(int N, const int *__restrict__ mask, const R *__restrict__ learned, const R *__restrict__ plain,
                            float4 *__restrict__ d) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float v[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const size_t i = (size_t)k * N + n;
    v[k] = mask[n] ? 0.0f : (float)(learned[i] - plain[i]);
  }
  d[n] = make_float4(v[0], v[1], v[2], 0.0f);
}
// One sweep; the neighbours come from SELL-64 slices (one coalesced index load per slot for the
// whole wavefront, a uniform loop to the slice width; padding slots hold the node itself).
__global__ void k_corr_smooth(int N, const int *__restrict__ mask, const int *__restrict__ adj_off,
                              const int *__restrict__ sell_off, const int *__restrict__ sell_col,
                              const float4 *__restrict__ d, float4 *__restrict__ dn) {
  // XCD-contiguous blocks (block i runs on XCD i % 8): a node's neighbours, mostly in the same or
  // the adjacent blocks in Morton order, stay in that XCD's L2
  const int per = (gridDim.x + 7) >> 3, xcd = blockIdx.x & 7, full = gridDim.x - 8 * (per - 1);
  const int blk = (xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1)) + (blockIdx.x >> 3);
  const int n = blk * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int e0 = adj_off[n], e1 = adj_off[n + 1];
  const int g0 = sell_off[n >> 6], width = sell_off[(n >> 6) + 1] - g0, lane = n & 63;
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
#pragma unroll 2
  for (int j = 0; j < width; ++j) {
    const int m = sell_col[(size_t)(g0 + j) * 64 + lane];
    const float4 v = d[m];
    if (m != n) {  // neighbours ascending as in the adjacency list; padding is the node itself
      a0 += v.x;
      a1 += v.y;
      a2 += v.z;
    }
  }
  const float ic = 1.0f / (float)(e1 - e0 > 0 ? e1 - e0 : 1);
  const float4 s = d[n];
  dn[n] = mask[n] ? make_float4(0.0f, 0.0f, 0.0f, 0.0f)
                  : make_float4(0.5f * s.x + 0.5f * (a0 * ic), 0.5f * s.y + 0.5f * (a1 * ic), 0.5f * s.z + 0.5f * (a2 * ic),
                                0.0f);
}
// Candidate ring.b[1] = ring.b[0] + beta d, with beta from the argument or (beta < 0) the
// control block's chosen step.
template <typename R>
__global__ void k_corr_apply(int N, const Ctrl *__restrict__ ctrl, float beta, const R *__restrict__ plain,
                             const float4 *__restrict__ d, R *__restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N || ctrl->done) return;
  const R b = (R)(beta < 0.0f ? ctrl->ws_beta : beta);
  const float4 v = d[n];
  const float dv[3] = {v.x, v.y, v.z};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const size_t i = (size_t)k * N + n;
    out[i] = plain[i] + b * (R)dv[k];
  }
}

// Warm-start selection around gradient passes (k_tile + k_node<R, true>, no Hessian refresh or
// energy):  stage 0 starts, 1 records |g(plain)|^2, 2 records |g(candidate beta)|^2 and keeps the
// best beta, 3 starts the descent from the best candidate iff |g| <= |g(plain)| (the plain norm
// stays the rel-grad denominator).
__global__ void __launch_bounds__(1024) k_ws_stage(Ctrl *__restrict__ ctrl, const double *__restrict__ npart, int nnb,
                                                   int stage, float beta) {
  __shared__ double lg[1024];
  if (ctrl->done) return;
  const int tid = threadIdx.x;
  double c = 0.0;
  if (stage == 1 || stage == 2)
    for (int i = tid; i < nnb; i += 1024) c += npart[2 * i + 1];
  lg[tid] = c;
  __syncthreads();
  for (int st = 512; st > 0; st >>= 1) {
    if (tid < st) lg[tid] += lg[tid + st];
    __syncthreads();
  }
  if (tid == 0) {
    Ctrl C = *ctrl;
    if (stage == 0) {
      C.hd_now = 0;
      C.check_now = 0;
      C.cur = 0;
      C.g_learned = -1.0;
      C.ws_beta = 0.0f;
    } else if (stage == 1) {
      C.g_plain = lg[0];
      C.cur = 1;
    } else if (stage == 2) {
      if (C.g_learned < 0.0 || lg[0] < C.g_learned) {
        C.g_learned = lg[0];
        C.ws_beta = beta;
      }
    } else {
      const bool use = C.g_learned >= 0.0 && C.g_learned <= C.g_plain;
      C.cur = use ? 1 : 0;
      C.nxt = use ? 0 : 1;
      C.used_learned = use ? 1 : 0;
      C.g0 = sqrt(C.g_plain);
      C.hd_now = 1;
      C.check_now = 1;
    }
    *ctrl = C;
  }
}

// Appends (features, x_new - y) of free nodes: sample base + free_rank[n].
template <typename R>
__global__ void k_collect(DevParams p, const R *__restrict__ xnew, const float *__restrict__ feat, const R *__restrict__ y,
                          const int *__restrict__ free_rank, long base, int stride, int phase,
                          float *__restrict__ ds_x, float *__restrict__ ds_y) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = p.nn;
  if (n >= N) return;
  const int r = free_rank[n];
  if (r < 0 || r % stride != phase) return;
  const size_t j = (size_t)(base + r / stride);
#pragma unroll
  for (int k = 0; k < 18; ++k) ds_x[j * 18 + k] = feat[(size_t)n * 18 + k];
#pragma unroll
  for (int d = 0; d < 3; ++d) ds_y[j * 3 + d] = (float)(xnew[(size_t)d * N + n] - y[(size_t)d * N + n]);
}

// ---------------------------------------------------------------------------------------------
// Training: one 256-thread workgroup per 32 samples, FP32 on the vector units with the BF16
// roundings of the inference path (inputs, weights and hidden activations), straight-through
// gradients for the roundings.  Loss = mean over (sample, component) of (out - t / scale)^2 with
// the mean taken over the whole batch.  Each workgroup writes its gradient partial (packed like
// the parameters) and its sum of squared errors.
// ---------------------------------------------------------------------------------------------
constexpr int TS = 32;  // samples per training workgroup
struct TrainLds {
  float W[9536];  // BF16-rounded weights, layers packed without biases
  float h0[TS][18];
  float z[3][TS][64], h[3][TS][64];
  float dA[TS][64], dB[TS][64];
  float d4[TS][3], t[TS][3];
  double red[4];
};

__global__ void __launch_bounds__(256) k_mlp_train(MlpDev net, const float *__restrict__ ds_x,
                                                   const float *__restrict__ ds_y, const int *__restrict__ idx, int n,
                                                   int batch, float *__restrict__ grad_part,
                                                   float *__restrict__ loss_part) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  TrainLds &S = *(TrainLds *)smem_raw;
  const int tid = threadIdx.x, w = blockIdx.x;
  const float *P = net.params;
  const int wofs[4] = {0, 1152, 5248, 9344};  // layer weight offsets in S.W
  for (int e = tid; e < 9536; e += 256) {
    int l = 3;
    while (l > 0 && e < wofs[l]) --l;
    S.W[e] = bfround(P[kLayers[l].pw + (e - wofs[l])]);
  }
  for (int e = tid; e < TS * 18; e += 256) {
    const int s = e / 18, k = e % 18, j = w * TS + s;
    const float f = j < n ? ds_x[(size_t)idx[j] * 18 + k] * net.inv_s[k / 3] : 0.0f;
    S.h0[s][k] = bfround(f);
  }
  for (int e = tid; e < TS * 3; e += 256) {
    const int s = e / 3, i = e % 3, j = w * TS + s;
    S.t[s][i] = j < n ? ds_y[(size_t)idx[j] * 3 + i] * net.inv_s[6] : 0.0f;
  }
  __syncthreads();
  // forward, hidden layers
  const int o = tid & 63, sg = tid >> 6;
  for (int l = 0; l < 3; ++l) {
    const int K = l == 0 ? 18 : 64;
    const float *Wl = S.W + wofs[l];
    const float bias = P[kLayers[l].pb + o];
    for (int jj = 0; jj < TS / 4; ++jj) {
      const int s = sg + 4 * jj;
      const float *hin = l == 0 ? S.h0[s] : S.h[l - 1][s];
      float acc = bias;
      for (int k = 0; k < K; ++k) acc = fmaf(Wl[k * 64 + o], hin[k], acc);
      S.z[l][s][o] = acc;
      S.h[l][s][o] = bfround(silu(acc));
    }
    __syncthreads();
  }
  // output layer, loss, output delta
  double se = 0.0;
  if (tid < TS * 3) {
    const int s = tid / 3, i = tid % 3, j = w * TS + s;
    float acc = P[9728 + i];
    for (int k = 0; k < 64; ++k) acc = fmaf(S.W[9344 + k * 3 + i], S.h[2][s][k], acc);
    const float diff = acc - S.t[s][i];
    const bool valid = j < n;
    S.d4[s][i] = valid ? (2.0f / (3.0f * (float)batch)) * diff : 0.0f;
    se = valid ? (double)diff * (double)diff : 0.0;
  }
  // block sum of squared errors (waves 0 and 1 hold the 96 values)
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) se += __shfl_down(se, off, 64);
  if ((tid & 63) == 0) S.red[tid >> 6] = se;
  __syncthreads();
  float *gp = grad_part + (size_t)w * MLP_NPARAM;
  if (tid == 0) loss_part[w] = (float)(S.red[0] + S.red[1] + S.red[2] + S.red[3]);
  // layer 4 gradients and delta into layer 3
  if (tid < 192) {
    const int k = tid / 3, i = tid % 3;
    float a = 0.0f;
    for (int s = 0; s < TS; ++s) a = fmaf(S.h[2][s][k], S.d4[s][i], a);
    gp[9536 + k * 3 + i] = a;
  } else if (tid < 195) {
    const int i = tid - 192;
    float a = 0.0f;
    for (int s = 0; s < TS; ++s) a += S.d4[s][i];
    gp[9728 + i] = a;
  }
  for (int jj = 0; jj < TS / 4; ++jj) {
    const int s = sg + 4 * jj, k = o;
    float a = 0.0f;
    for (int i = 0; i < 3; ++i) a = fmaf(S.W[9344 + k * 3 + i], S.d4[s][i], a);
    S.dA[s][k] = a * dsilu(S.z[2][s][k]);
  }
  __syncthreads();
  // hidden layers 3, 2, 1 (l = 2, 1, 0): dW = h_{l-1}^T dz, db = sum dz, dh_{l-1} = W dz
  float(*dz)[64] = S.dA;
  float(*dn)[64] = S.dB;
  for (int l = 2; l >= 0; --l) {
    const int K = l == 0 ? 18 : 64;
    const int pw = kLayers[l].pw, pb = kLayers[l].pb;
    for (int e = tid; e < K * 64; e += 256) {
      const int k = e / 64, oo = e % 64;
      float a = 0.0f;
      for (int s = 0; s < TS; ++s) a = fmaf(l == 0 ? S.h0[s][k] : S.h[l - 1][s][k], dz[s][oo], a);
      gp[pw + k * 64 + oo] = a;
    }
    if (tid < 64) {
      float a = 0.0f;
      for (int s = 0; s < TS; ++s) a += dz[s][tid];
      gp[pb + tid] = a;
    }
    if (l > 0) {
      const float *Wl = S.W + wofs[l];
      for (int jj = 0; jj < TS / 4; ++jj) {
        const int s = sg + 4 * jj, k = o;
        float a = 0.0f;
        for (int oo = 0; oo < 64; ++oo) a = fmaf(Wl[k * 64 + oo], dz[s][oo], a);
        dn[s][k] = a * dsilu(S.z[l - 1][s][k]);
      }
      __syncthreads();
      float(*tmp)[64] = dz;
      dz = dn;
      dn = tmp;
    }
  }
}

// grad[p] = sum over workgroups (in order) of the partials; loss_acc += sum of squared errors.
__global__ void k_mlp_reduce(const float *__restrict__ grad_part, const float *__restrict__ loss_part, int nwg,
                             float *__restrict__ grad, double *__restrict__ loss_acc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < MLP_NPARAM) {
    float a = 0.0f;
    for (int w = 0; w < nwg; ++w) a += grad_part[(size_t)w * MLP_NPARAM + p];
    grad[p] = a;
  }
  if (p == 0) {
    double l = 0.0;
    for (int w = 0; w < nwg; ++w) l += (double)loss_part[w];
    *loss_acc += l;
  }
}

// Adam (beta1 = 0.9, beta2 = 0.999, eps = 1e-8); b1t, b2t = beta^t.
__global__ void k_mlp_adam(float *__restrict__ params, float *__restrict__ m1, float *__restrict__ m2,
                           const float *__restrict__ grad, float lr, float b1t, float b2t) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  const float g = grad[p];
  const float a = 0.9f * m1[p] + 0.1f * g;
  const float b = 0.999f * m2[p] + 0.001f * g * g;
  m1[p] = a;
  m2[p] = b;
  const float mh = a / (1.0f - b1t), vh = b / (1.0f - b2t);
  params[p] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

// ---- launch wrappers ----
static inline unsigned nb(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename R>
void LearnLaunch<R>::features(const DevMesh<R> &m, const DevParams &p, const int *adj_off, const int *adj, float *feat,
                              R *y, hipStream_t st) {
  hipLaunchKernelGGL((k_features<R>), dim3(nb(p.nn, 256)), dim3(256), 0, st, p, (const R *)m.xpos, (const R *)m.v,
                     m.rest, m.mask, adj_off, adj, feat, y);
}
template <typename R>
void LearnLaunch<R>::guess(const DevMesh<R> &m, const DevParams &p, const MlpDev &net, const float *feat, const R *y,
                           hipStream_t st) {
  const long tiles = (p.nn + 15) / 16;
  const unsigned grid = (unsigned)std::min<long>((tiles + 3) / 4, 2048);
  hipLaunchKernelGGL((k_guess<R>), dim3(grid), dim3(256), 0, st, p, net, feat, y, m.mask, (const R *)m.ring.b[0],
                     m.ring.b[1], 1.0f / net.inv_s[6]);
}
template <typename R>
void LearnLaunch<R>::select(const DevMesh<R> &m, const TileDev &td, const DevParams &p, const int *adj_off,
                            const int *adj, const int *sell_off, const int *sell_col, int iters, const float *betas,
                            int nbeta, float *sm, hipStream_t st) {
  const unsigned g = nb(p.nn, 256);
  const int nnb = (int)nb(p.nn, 256);
  float4 *sm4[2] = {(float4 *)sm, (float4 *)sm + p.nn};
  hipLaunchKernelGGL((k_corr_diff<R>), dim3(g), dim3(256), 0, st, p.nn, m.mask, (const R *)m.ring.b[1],
                     (const R *)m.ring.b[0], sm4[0]);
  for (int k = 0; k < iters; ++k)
    hipLaunchKernelGGL(k_corr_smooth, dim3(g), dim3(256), 0, st, p.nn, m.mask, adj_off, sell_off, sell_col,
                       (const float4 *)sm4[k & 1], sm4[(k + 1) & 1]);
  const float4 *d = sm4[iters & 1];
  hipLaunchKernelGGL(k_ws_stage, dim3(1), dim3(1024), 0, st, m.ctrl, td.npart, nnb, 0, 0.0f);
  Launch<R>::tile_pass(m, td, p, st);
  Launch<R>::grad_pass(m, td, p, st);
  hipLaunchKernelGGL(k_ws_stage, dim3(1), dim3(1024), 0, st, m.ctrl, td.npart, nnb, 1, 0.0f);
  for (int j = 0; j < nbeta; ++j) {
    hipLaunchKernelGGL((k_corr_apply<R>), dim3(g), dim3(256), 0, st, p.nn, (const Ctrl *)m.ctrl, betas[j],
                       (const R *)m.ring.b[0], d, m.ring.b[1]);
    Launch<R>::tile_pass(m, td, p, st);
    Launch<R>::grad_pass(m, td, p, st);
    hipLaunchKernelGGL(k_ws_stage, dim3(1), dim3(1024), 0, st, m.ctrl, td.npart, nnb, 2, betas[j]);
  }
  hipLaunchKernelGGL((k_corr_apply<R>), dim3(g), dim3(256), 0, st, p.nn, (const Ctrl *)m.ctrl, -1.0f,
                     (const R *)m.ring.b[0], d, m.ring.b[1]);
  hipLaunchKernelGGL(k_ws_stage, dim3(1), dim3(1024), 0, st, m.ctrl, td.npart, nnb, 3, 0.0f);
}
template <typename R>
void LearnLaunch<R>::collect(const DevMesh<R> &m, const DevParams &p, const float *feat, const R *y,
                             const int *free_rank, long base, int stride, int phase, float *ds_x, float *ds_y,
                             hipStream_t st) {
  hipLaunchKernelGGL((k_collect<R>), dim3(nb(p.nn, 256)), dim3(256), 0, st, p, (const R *)m.xpos, feat, y, free_rank,
                     base, stride, phase, ds_x, ds_y);
}
template struct LearnLaunch<float>;
template struct LearnLaunch<double>;

void mlp_arrange(const float *params, unsigned short *wA, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_arrange, dim3(nb(MLP_WA, 256)), dim3(256), 0, st, params, wA);
}
void mlp_forward(const MlpDev &net, const float *feat, int n, float *out, hipStream_t st) {
  if (n <= 0) return;
  const long tiles = (n + 15) / 16;
  const unsigned grid = (unsigned)std::min<long>((tiles + 3) / 4, 2048);
  hipLaunchKernelGGL(k_mlp_forward, dim3(grid), dim3(256), 0, st, net, feat, n, out);
}
void mlp_train_partials(const MlpDev &net, const float *ds_x, const float *ds_y, const int *idx, int n, int batch,
                        float *grad_part, float *loss_part, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mlp_train, dim3(nb(n, TS)), dim3(256), sizeof(TrainLds), st, net, ds_x, ds_y, idx, n, batch,
                     grad_part, loss_part);
}
void mlp_reduce(const float *grad_part, const float *loss_part, int nwg, float *grad, double *loss_acc, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_reduce, dim3(nb(MLP_NPARAM, 256)), dim3(256), 0, st, grad_part, loss_part, nwg, grad,
                     loss_acc);
}
void mlp_adam(float *params, float *m1, float *m2, const float *grad, float lr, float b1t, float b2t, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_adam, dim3(nb(MLP_NPARAM, 256)), dim3(256), 0, st, params, m1, m2, grad, lr, b1t, b2t);
}

}  // namespace pbs
