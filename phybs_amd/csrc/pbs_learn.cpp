// pbs_learn.cpp -- C ABI of the learned warm start (include/phybs.h, "learned warm start"): model
// storage, sample collection from pbs_step, training on the GPU, and the step hooks.  No
// reference counterpart.
#include "pbs_learn.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "pbs_solver.h"

namespace pbs {

template <typename T> static int lalloc(LearnState *L, T **p, size_t n) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(PBS_ERR_HIP, std::string("hipMalloc(") + std::to_string(n * sizeof(T)) + " B): " + hipGetErrorString(e));
  }
  L->allocs.push_back(q);
  *p = (T *)q;
  return PBS_OK;
}
static void lfree(LearnState *L, void *p) {
  if (!p) return;
  auto it = std::find(L->allocs.begin(), L->allocs.end(), p);
  if (it != L->allocs.end()) L->allocs.erase(it);
  (void)hipFree(p);
}

void learn_release(LearnState *L) {
  if (!L) return;
  for (void *q : L->allocs) (void)hipFree(q);
  delete L;
}

MlpDev mlp_dev(const LearnState &L) {
  MlpDev d{};
  d.params = L.params;
  d.wA = L.wA;
  for (int i = 0; i < MLP_NSCALE; ++i) d.inv_s[i] = 1.0f / L.scales[i];
  return d;
}

// Creates the state with the network buffers, node adjacency and free-node ranks (solver order).
static int learn_get(pbs_solver *s, LearnState **out) {
  if (s->learn) {
    *out = s->learn;
    return PBS_OK;
  }
  LearnState *L = new LearnState();
  const size_t N = s->nn;
  const size_t rsz = s->p.precision == PBS_FP64 ? 8 : 4;
  int rc = 0;
  rc |= lalloc(L, &L->params, MLP_NPARAM);
  rc |= lalloc(L, &L->m1, MLP_NPARAM);
  rc |= lalloc(L, &L->m2, MLP_NPARAM);
  rc |= lalloc(L, &L->grad, MLP_NPARAM);
  rc |= lalloc(L, &L->wA, MLP_WA);
  rc |= lalloc(L, &L->feat, 18 * N);
  unsigned char *yb = nullptr;
  rc |= lalloc(L, &yb, 3 * N * rsz);
  L->y = yb;
  rc |= lalloc(L, &L->loss_acc, 1);
  // node adjacency from the tets (solver numbering), neighbours ascending, without the node
  const TilePlan &P = s->plan;
  std::vector<int> off(N + 1, 0), adj;
  adj.reserve(14 * N);
  std::vector<int> nb;
  for (size_t n = 0; n < N; ++n) {
    nb.clear();
    for (int e = P.inc_off[n]; e < P.inc_off[n + 1]; ++e) {
      const int t = P.inc[e] >> 2;
      for (int j = 0; j < 4; ++j) {
        const int m = P.tets_new[4 * (size_t)t + j];
        if (m != (int)n) nb.push_back(m);
      }
    }
    std::sort(nb.begin(), nb.end());
    nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
    adj.insert(adj.end(), nb.begin(), nb.end());
    off[n + 1] = (int)adj.size();
  }
  rc |= lalloc(L, &L->adj_off, N + 1);
  rc |= lalloc(L, &L->adj, std::max<size_t>(adj.size(), 1));
  rc |= lalloc(L, &L->free_rank, N);
  rc |= lalloc(L, &L->smooth, 8 * N);
  // the same adjacency as SELL-64 slices for the smoothing sweeps: slot j of row l of slice s at
  // (sell_off[s] + j) * 64 + l, rows padded with the node itself to the slice's longest row
  const size_t nsl = (N + 63) / 64;
  std::vector<int> soff(nsl + 1, 0);
  for (size_t sl = 0; sl < nsl; ++sl) {
    int w = 0;
    for (size_t n = 64 * sl; n < std::min(N, 64 * sl + 64); ++n) w = std::max(w, off[n + 1] - off[n]);
    soff[sl + 1] = soff[sl] + w;
  }
  std::vector<int> scol(64 * (size_t)soff[nsl]);
  for (size_t n = 0; n < N; ++n) {
    const size_t g0 = soff[n >> 6], w = soff[(n >> 6) + 1] - g0, lane = n & 63;
    for (size_t j = 0; j < w; ++j)
      scol[(g0 + j) * 64 + lane] = (int)j < off[n + 1] - off[n] ? adj[off[n] + j] : (int)n;
  }
  rc |= lalloc(L, &L->sell_off, nsl + 1);
  rc |= lalloc(L, &L->sell_col, std::max<size_t>(scol.size(), 1));
  if (rc) {
    learn_release(L);
    return PBS_ERR_HIP;
  }
  // free-node ranks in the caller's node order (samples of a step are stored in that order),
  // indexed by solver node; the device mask is in solver order
  std::vector<int> mask(N), rank(N);
  hipError_t e = hipMemcpyAsync(mask.data(), s->p.precision == PBS_FP64 ? s->m64.mask : s->m32.mask, sizeof(int) * N,
                                hipMemcpyDeviceToHost, s->st);
  if (e == hipSuccess) e = hipStreamSynchronize(s->st);
  long nf = 0;
  for (size_t o = 0; o < N; ++o) {
    const int i = P.iperm[o];
    rank[i] = mask[i] ? -1 : (int)nf++;
  }
  L->nfree = nf;
  if (e == hipSuccess) e = hipMemcpyAsync(L->adj_off, off.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice, s->st);
  if (e == hipSuccess && !adj.empty())
    e = hipMemcpyAsync(L->adj, adj.data(), sizeof(int) * adj.size(), hipMemcpyHostToDevice, s->st);
  if (e == hipSuccess) e = hipMemcpyAsync(L->free_rank, rank.data(), sizeof(int) * N, hipMemcpyHostToDevice, s->st);
  if (e == hipSuccess) e = hipMemcpyAsync(L->sell_off, soff.data(), sizeof(int) * soff.size(), hipMemcpyHostToDevice, s->st);
  if (e == hipSuccess && !scol.empty())
    e = hipMemcpyAsync(L->sell_col, scol.data(), sizeof(int) * scol.size(), hipMemcpyHostToDevice, s->st);
  if (e == hipSuccess) e = hipMemsetAsync(L->m1, 0, sizeof(float) * MLP_NPARAM, s->st);
  if (e == hipSuccess) e = hipMemsetAsync(L->m2, 0, sizeof(float) * MLP_NPARAM, s->st);
  if (e == hipSuccess) e = hipStreamSynchronize(s->st);
  if (e != hipSuccess) {
    learn_release(L);
    return set_error(PBS_ERR_HIP, std::string("learned warm start setup: ") + hipGetErrorString(e));
  }
  s->learn = L;
  *out = L;
  return PBS_OK;
}

static int upload_model(pbs_solver *s, LearnState *L, const float *params) {
  HIPCHK(hipMemcpyAsync(L->params, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, s->st));
  HIPCHK(hipMemsetAsync(L->m1, 0, sizeof(float) * MLP_NPARAM, s->st));
  HIPCHK(hipMemsetAsync(L->m2, 0, sizeof(float) * MLP_NPARAM, s->st));
  mlp_arrange(L->params, L->wA, s->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->st));
  L->have_model = true;
  L->adam_t = 0;
  return PBS_OK;
}

static int io_reserve(LearnState *L, long n) {
  if (L->io_cap >= n) return PBS_OK;
  lfree(L, L->io);
  L->io = nullptr;
  L->io_cap = 0;
  int rc = lalloc(L, &L->io, (size_t)n);
  if (!rc) L->io_cap = n;
  return rc;
}

// ---- step hooks ----
template <typename R> int learn_before_descent(pbs_solver *s) {
  LearnState *L = s->learn;
  if (!L || !(L->collecting || (L->enabled && L->have_model))) return PBS_OK;
  DevMesh<R> &m = mesh<R>(s);
  LearnLaunch<R>::features(m, s->dp, L->adj_off, L->adj, L->feat, (R *)L->y, s->st);
  if (L->enabled && L->have_model) {
    LearnLaunch<R>::guess(m, s->dp, mlp_dev(*L), L->feat, (const R *)L->y, s->st);
    LearnLaunch<R>::select(m, s->td, s->dp, L->adj_off, L->adj, L->sell_off, L->sell_col, L->smooth_iters, L->betas.data(), (int)L->betas.size(),
                           L->smooth, s->st);
  }
  HIPCHK(hipGetLastError());
  return PBS_OK;
}
template <typename R> int learn_after_step(pbs_solver *s) {
  LearnState *L = s->learn;
  if (!L || !L->collecting) return PBS_OK;
  const int phase = (int)(L->steps_seen % L->stride);
  const long take = L->nfree > phase ? (L->nfree - phase + L->stride - 1) / L->stride : 0;
  if (L->count + take > L->cap) {
    L->collecting = false;  // full
    return PBS_OK;
  }
  LearnLaunch<R>::collect(mesh<R>(s), s->dp, L->feat, (const R *)L->y, L->free_rank, L->count, L->stride, phase,
                          L->ds_x, L->ds_y, s->st);
  HIPCHK(hipGetLastError());
  L->count += take;
  L->steps_seen++;
  return PBS_OK;
}
template int learn_before_descent<float>(pbs_solver *);
template int learn_before_descent<double>(pbs_solver *);
template int learn_after_step<float>(pbs_solver *);
template int learn_after_step<double>(pbs_solver *);

}  // namespace pbs

using namespace pbs;

#define LEARN_GET(s, L)                                       \
  if (!(s)) return set_error(PBS_ERR_ARG, "null handle");     \
  if ((s)->ninst != 1) return set_error(PBS_ERR_STATE, "the learned warm start needs a single-mesh solver"); \
  LearnState *L = nullptr;                                    \
  {                                                           \
    int rc_ = learn_get((s), &L);                             \
    if (rc_) return rc_;                                      \
  }

extern "C" int pbs_learn_set_model(pbs_solver *s, const float *params, const float *scales) {
  LEARN_GET(s, L);
  if (!params || !scales) return set_error(PBS_ERR_ARG, "null model");
  for (int i = 0; i < MLP_NSCALE; ++i)
    if (!(scales[i] > 0) || !std::isfinite(scales[i])) return set_error(PBS_ERR_ARG, "scales must be positive");
  std::memcpy(L->scales, scales, sizeof(L->scales));
  return upload_model(s, L, params);
}

extern "C" int pbs_learn_get_model(pbs_solver *s, float *params, float *scales) {
  LEARN_GET(s, L);
  if (!L->have_model) return set_error(PBS_ERR_STATE, "no model");
  if (params) {
    HIPCHK(hipMemcpyAsync(params, L->params, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
  }
  if (scales) std::memcpy(scales, L->scales, sizeof(L->scales));
  return PBS_OK;
}

extern "C" int pbs_learn_init_model(pbs_solver *s, uint64_t seed) {
  LEARN_GET(s, L);
  // Glorot-uniform weights, zero biases
  std::mt19937_64 rng(seed);
  std::vector<float> p;
  p.reserve(MLP_NPARAM);
  const int sizes[4][2] = {{18, 64}, {64, 64}, {64, 64}, {64, 3}};
  for (auto &sz : sizes) {
    const double lim = std::sqrt(6.0 / (sz[0] + sz[1]));
    std::uniform_real_distribution<double> U(-lim, lim);
    for (int i = 0; i < sz[0] * sz[1]; ++i) p.push_back((float)U(rng));
    for (int i = 0; i < sz[1]; ++i) p.push_back(0.0f);
  }
  return upload_model(s, L, p.data());
}

extern "C" int pbs_learn_enable(pbs_solver *s, int on) {
  LEARN_GET(s, L);
  if (on && !L->have_model) return set_error(PBS_ERR_STATE, "no model: set, initialise or train one first");
  L->enabled = on != 0;
  return PBS_OK;
}

extern "C" int pbs_learn_set_options(pbs_solver *s, int sweeps, int nbeta, const float *betas) {
  LEARN_GET(s, L);
  if (sweeps < 0 || sweeps > 10000) return set_error(PBS_ERR_ARG, "sweeps must be in [0, 10000]");
  if (nbeta < 1 || nbeta > 16 || !betas) return set_error(PBS_ERR_ARG, "need 1..16 steps");
  for (int j = 0; j < nbeta; ++j)
    if (!(betas[j] >= 0.0f) || !std::isfinite(betas[j])) return set_error(PBS_ERR_ARG, "steps must be >= 0");
  L->smooth_iters = sweeps;
  L->betas.assign(betas, betas + nbeta);
  return PBS_OK;
}

extern "C" int pbs_learn_collect(pbs_solver *s, int64_t capacity) { return pbs_learn_collect_every(s, capacity, 1); }

extern "C" int pbs_learn_collect_every(pbs_solver *s, int64_t capacity, int stride) {
  LEARN_GET(s, L);
  if (capacity < 0) return set_error(PBS_ERR_ARG, "capacity must be >= 0");
  if (stride < 1) return set_error(PBS_ERR_ARG, "stride must be >= 1");
  lfree(L, L->ds_x);
  lfree(L, L->ds_y);
  L->ds_x = L->ds_y = nullptr;
  L->cap = L->count = 0;
  L->stride = stride;
  L->steps_seen = 0;
  L->collecting = false;
  if (capacity == 0) return PBS_OK;
  int rc = lalloc(L, &L->ds_x, 18 * (size_t)capacity);
  if (!rc) rc = lalloc(L, &L->ds_y, 3 * (size_t)capacity);
  if (rc) return rc;
  L->cap = capacity;
  L->collecting = true;
  return PBS_OK;
}

extern "C" int64_t pbs_learn_samples(pbs_solver *s) {
  if (!s) return set_error(PBS_ERR_ARG, "null handle");
  return s->learn ? s->learn->count : 0;
}

extern "C" int pbs_learn_get_samples(pbs_solver *s, int64_t start, int64_t n, float *features, float *targets) {
  LEARN_GET(s, L);
  if (start < 0 || n < 0 || start + n > L->count) return set_error(PBS_ERR_ARG, "sample range out of bounds");
  if (features && n)
    HIPCHK(hipMemcpyAsync(features, L->ds_x + 18 * start, sizeof(float) * 18 * n, hipMemcpyDeviceToHost, s->st));
  if (targets && n)
    HIPCHK(hipMemcpyAsync(targets, L->ds_y + 3 * start, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return PBS_OK;
}

extern "C" int pbs_learn_add_samples(pbs_solver *s, int64_t n, const float *features, const float *targets) {
  LEARN_GET(s, L);
  if (n < 0 || (n > 0 && (!features || !targets))) return set_error(PBS_ERR_ARG, "pbs_learn_add_samples: bad arguments");
  if (L->count + n > L->cap) return set_error(PBS_ERR_ARG, "pbs_learn_add_samples: sample buffer too small");
  if (n) {
    HIPCHK(hipMemcpyAsync(L->ds_x + 18 * L->count, features, sizeof(float) * 18 * n, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemcpyAsync(L->ds_y + 3 * L->count, targets, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
  }
  L->count += n;
  return PBS_OK;
}

// Scales = RMS over the collected samples of each feature group and of the targets.
extern "C" int pbs_learn_fit_scales(pbs_solver *s) {
  LEARN_GET(s, L);
  if (L->count == 0) return set_error(PBS_ERR_STATE, "no samples collected");
  std::vector<float> x(18 * (size_t)L->count), y(3 * (size_t)L->count);
  HIPCHK(hipMemcpyAsync(x.data(), L->ds_x, sizeof(float) * x.size(), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipMemcpyAsync(y.data(), L->ds_y, sizeof(float) * y.size(), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  double acc[MLP_NSCALE] = {0};
  for (long j = 0; j < L->count; ++j) {
    for (int k = 0; k < 18; ++k) acc[k / 3] += (double)x[18 * j + k] * x[18 * j + k];
    for (int k = 0; k < 3; ++k) acc[6] += (double)y[3 * j + k] * y[3 * j + k];
  }
  for (int g = 0; g < MLP_NSCALE; ++g) {
    const double r = std::sqrt(acc[g] / (3.0 * (double)L->count));
    L->scales[g] = r > 0 && std::isfinite(r) ? (float)r : 1.0f;
  }
  return PBS_OK;
}

static int train_batch(pbs_solver *s, LearnState *L, const float *ds_x, const float *ds_y, const int *idx, int n,
                       float lr, bool update) {
  const int nwg = (n + 31) / 32;
  if (nwg > L->part_cap) {
    lfree(L, L->grad_part);
    lfree(L, L->loss_part);
    L->grad_part = L->loss_part = nullptr;
    L->part_cap = 0;
    int rc = lalloc(L, &L->grad_part, (size_t)nwg * MLP_NPARAM);
    if (!rc) rc = lalloc(L, &L->loss_part, (size_t)nwg);
    if (rc) return rc;
    L->part_cap = nwg;
  }
  const MlpDev net = mlp_dev(*L);
  mlp_train_partials(net, ds_x, ds_y, idx, n, n, L->grad_part, L->loss_part, s->st);
  mlp_reduce(L->grad_part, L->loss_part, nwg, L->grad, L->loss_acc, s->st);
  if (update) {
    L->adam_t++;
    const float b1t = (float)std::pow(0.9, (double)L->adam_t), b2t = (float)std::pow(0.999, (double)L->adam_t);
    mlp_adam(L->params, L->m1, L->m2, L->grad, lr, b1t, b2t, s->st);
    mlp_arrange(L->params, L->wA, s->st);
  }
  HIPCHK(hipGetLastError());
  return PBS_OK;
}

extern "C" int pbs_learn_train(pbs_solver *s, int epochs, int batch, float lr, uint64_t seed, float *loss_per_epoch) {
  LEARN_GET(s, L);
  if (!L->have_model) return set_error(PBS_ERR_STATE, "no model: initialise one first");
  if (L->count == 0) return set_error(PBS_ERR_STATE, "no samples collected");
  if (epochs < 0 || batch <= 0 || !(lr > 0)) return set_error(PBS_ERR_ARG, "bad training arguments");
  const long n = L->count;
  if (L->idx_cap < n) {
    lfree(L, L->idx);
    L->idx = nullptr;
    L->idx_cap = 0;
    int rc = lalloc(L, &L->idx, (size_t)n);
    if (rc) return rc;
    L->idx_cap = n;
  }
  std::mt19937_64 rng(seed);
  std::vector<int> perm(n);
  for (int ep = 0; ep < epochs; ++ep) {
    std::iota(perm.begin(), perm.end(), 0);
    std::shuffle(perm.begin(), perm.end(), rng);
    HIPCHK(hipMemcpyAsync(L->idx, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemsetAsync(L->loss_acc, 0, sizeof(double), s->st));
    for (long b0 = 0; b0 < n; b0 += batch) {
      const int nb = (int)std::min<long>(batch, n - b0);
      int rc = train_batch(s, L, L->ds_x, L->ds_y, L->idx + b0, nb, lr, true);
      if (rc) return rc;
    }
    double acc = 0.0;
    HIPCHK(hipMemcpyAsync(&acc, L->loss_acc, sizeof(double), hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    if (loss_per_epoch) loss_per_epoch[ep] = (float)(acc / (3.0 * (double)n));
  }
  HIPCHK(hipStreamSynchronize(s->st));
  return PBS_OK;
}

extern "C" int pbs_learn_forward(pbs_solver *s, int n, const float *features, float *out) {
  LEARN_GET(s, L);
  if (n < 0 || (n && (!features || !out))) return set_error(PBS_ERR_ARG, "bad arguments");
  if (!L->have_model) return set_error(PBS_ERR_STATE, "no model");
  if (n == 0) return PBS_OK;
  int rc = io_reserve(L, 21L * n);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(L->io, features, sizeof(float) * 18 * n, hipMemcpyHostToDevice, s->st));
  mlp_forward(mlp_dev(*L), L->io, n, L->io + 18L * n, s->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, L->io + 18L * n, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return PBS_OK;
}

extern "C" int pbs_learn_loss_grad(pbs_solver *s, int n, const float *features, const float *targets, double *loss,
                                   float *grad) {
  LEARN_GET(s, L);
  if (n <= 0 || !features || !targets) return set_error(PBS_ERR_ARG, "bad arguments");
  if (!L->have_model) return set_error(PBS_ERR_STATE, "no model");
  int rc = io_reserve(L, 22L * n);
  if (rc) return rc;
  float *dx = L->io, *dy = L->io + 18L * n;
  int *idx = (int *)(L->io + 21L * n);
  std::vector<int> id(n);
  std::iota(id.begin(), id.end(), 0);
  HIPCHK(hipMemcpyAsync(dx, features, sizeof(float) * 18 * n, hipMemcpyHostToDevice, s->st));
  HIPCHK(hipMemcpyAsync(dy, targets, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s->st));
  HIPCHK(hipMemcpyAsync(idx, id.data(), sizeof(int) * n, hipMemcpyHostToDevice, s->st));
  HIPCHK(hipMemsetAsync(L->loss_acc, 0, sizeof(double), s->st));
  rc = train_batch(s, L, dx, dy, idx, n, 0.0f, false);
  if (rc) return rc;
  double acc = 0.0;
  HIPCHK(hipMemcpyAsync(&acc, L->loss_acc, sizeof(double), hipMemcpyDeviceToHost, s->st));
  if (grad) HIPCHK(hipMemcpyAsync(grad, L->grad, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  if (loss) *loss = acc / (3.0 * (double)n);
  return PBS_OK;
}

// Features of the current state, in the caller's node order (N x 18).
template <typename R> static int features_impl(pbs_solver *s, LearnState *L, float *out) {
  LearnLaunch<R>::features(mesh<R>(s), s->dp, L->adj_off, L->adj, L->feat, (R *)L->y, s->st);
  HIPCHK(hipGetLastError());
  std::vector<float> f(18 * (size_t)s->nn);
  HIPCHK(hipMemcpyAsync(f.data(), L->feat, sizeof(float) * f.size(), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  const std::vector<int> &perm = s->plan.perm;
  for (int i = 0; i < s->nn; ++i) std::memcpy(out + 18 * (size_t)perm[i], &f[18 * (size_t)i], 18 * sizeof(float));
  return PBS_OK;
}
extern "C" int pbs_learn_features(pbs_solver *s, float *features) {
  LEARN_GET(s, L);
  if (!features) return set_error(PBS_ERR_ARG, "null output");
  return s->p.precision == PBS_FP64 ? features_impl<double>(s, L, features) : features_impl<float>(s, L, features);
}

// The two initial guesses of the next step (caller order, AoS doubles); the solver state is not
// changed (the step-start kernel only writes the descent's scratch buffers).
template <typename R> static int guess_impl(pbs_solver *s, LearnState *L, double *learned, double *plain) {
  DevMesh<R> &m = mesh<R>(s);
  Launch<R>::step_init(m, s->td, s->dp, s->bs, s->st);
  LearnLaunch<R>::features(m, s->dp, L->adj_off, L->adj, L->feat, (R *)L->y, s->st);
  LearnLaunch<R>::guess(m, s->dp, mlp_dev(*L), L->feat, (const R *)L->y, s->st);
  LearnLaunch<R>::select(m, s->td, s->dp, L->adj_off, L->adj, L->sell_off, L->sell_col, L->smooth_iters, L->betas.data(), (int)L->betas.size(),
                         L->smooth, s->st);
  HIPCHK(hipGetLastError());
  for (int which = 0; which < 2; ++which) {
    double *dst = which ? learned : plain;
    if (!dst) continue;
    Launch<R>::planes_to_aos64(m.ring.b[which], s->stage64, m.perm, s->nn, s->st);
    HIPCHK(hipMemcpyAsync(dst, s->stage64, sizeof(double) * 3 * (size_t)s->nn, hipMemcpyDeviceToHost, s->st));
  }
  HIPCHK(hipStreamSynchronize(s->st));
  return PBS_OK;
}
extern "C" int pbs_learn_guess(pbs_solver *s, double *learned, double *plain) {
  LEARN_GET(s, L);
  if (!L->have_model) return set_error(PBS_ERR_STATE, "no model");
  return s->p.precision == PBS_FP64 ? guess_impl<double>(s, L, learned, plain) : guess_impl<float>(s, L, learned, plain);
}
