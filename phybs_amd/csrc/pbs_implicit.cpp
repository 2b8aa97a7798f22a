// pbs_implicit.cpp -- C ABI of the backward-Euler step (include/phybs.h, "backward Euler"):
// matrix layout construction, the conjugate-gradient driver (HIP graphs of CG iterations
// replayed until the device-side test stops them) and the stiffness-matrix export.
#include "pbs_implicit.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "pbs_solver.h"

namespace pbs {

void implicit_params_changed(ImplicitState *S) {
  if (S) S->pd_built = false;
}

void implicit_release(ImplicitState *S) {
  if (!S) return;
  for (CgGraphs *G : {&S->cg_ie, &S->cg_pd})
    if (G->graph) (void)hipGraphExecDestroy(G->graph);
  for (auto e : S->ev) (void)hipEventDestroy(e);
  for (void *q : S->allocs) (void)hipFree(q);
  delete S;
}

}  // namespace pbs

namespace {

constexpr int kGraphIters = 16;  // CG iterations per graph replay

template <typename T> int ialloc(ImplicitState *S, T **p, size_t n) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(PBS_ERR_HIP, std::string("hipMalloc(") + std::to_string(n * sizeof(T)) + " B): " + hipGetErrorString(e));
  }
  S->allocs.push_back(q);
  *p = (T *)q;
  return PBS_OK;
}

// Block rows of K in solver order: a node's columns are the nodes sharing a tet with it,
// ascending, itself included (the block pattern of allocateGlobalStiffnessMatrix,
// IsotropicMaterial.cpp:554-644), laid out as SELL-64.
int implicit_get(pbs_solver *s, ImplicitState **out) {
  if (s->implicit) {
    *out = s->implicit;
    return PBS_OK;
  }
  if (s->ninst != 1) return set_error(PBS_ERR_STATE, "backward-Euler step: not available on a batched solver");
  const int N = s->nn;
  const TilePlan &P = s->plan;
  std::vector<int> row_off(N + 1, 0), cols, tmp;
  cols.reserve((size_t)N * 16);
  for (int n = 0; n < N; ++n) {
    tmp.clear();
    for (int e = P.inc_off[n]; e < P.inc_off[n + 1]; ++e) {
      const int t = P.inc[e] >> 2;
      for (int c = 0; c < 4; ++c) tmp.push_back(P.tets_new[4 * (size_t)t + c]);
    }
    tmp.push_back(n);
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    cols.insert(cols.end(), tmp.begin(), tmp.end());
    row_off[n + 1] = (int)cols.size();
  }
  ImplicitState *S = new ImplicitState();
  S->nslice = (N + 63) / 64;
  S->slice_off_h.assign(S->nslice + 1, 0);
  S->row_len_h.resize(N);
  for (int n = 0; n < N; ++n) S->row_len_h[n] = row_off[n + 1] - row_off[n];
  for (int sl = 0; sl < S->nslice; ++sl) {
    int w = 0;
    for (int n = 64 * sl; n < std::min(N, 64 * sl + 64); ++n) w = std::max(w, S->row_len_h[n]);
    S->slice_off_h[sl + 1] = S->slice_off_h[sl] + w;
  }
  S->slots = 64L * S->slice_off_h[S->nslice];
  for (int sl = 0; sl < S->nslice; ++sl)
    S->max_width = std::max(S->max_width, S->slice_off_h[sl + 1] - S->slice_off_h[sl]);
  S->bcol_h.assign(S->slots, 0);
  std::vector<int> diag(N);
  for (int n = 0; n < N; ++n) {
    const long g0 = S->slice_off_h[n >> 6], w = S->slice_off_h[(n >> 6) + 1] - g0;
    const int lane = n & 63;
    for (long j = 0; j < w; ++j) {
      const bool real = j < S->row_len_h[n];
      const int c = real ? cols[row_off[n] + j] : n;
      S->bcol_h[(g0 + j) * 64 + lane] = c;
      if (real && c == n) diag[n] = (int)((g0 + j) * 64 + lane);
    }
  }
  const size_t rsz = s->p.precision == PBS_FP64 ? 8 : 4;
  const size_t n3 = 3 * (size_t)N;
  const int nb = (3 * N + 255) / 256;  // partial sums: up to one per 256 vector entries
  IeDev &d = S->dev;
  d.n = N;
  d.nslice = S->nslice;
  int *so = nullptr, *rl = nullptr, *bc = nullptr, *dsl = nullptr;
  char *A = nullptr;
  int rc = ialloc(S, &so, S->nslice + 1);
  if (!rc) rc = ialloc(S, &rl, N);
  if (!rc) rc = ialloc(S, &bc, S->slots);
  if (!rc) rc = ialloc(S, &dsl, N);
  if (!rc) rc = ialloc(S, &A, 9 * rsz * S->slots);
  double **vec[7] = {&d.b, &d.x, &d.r, &d.z, &d.p, &d.Ap, &d.dinv};
  for (int i = 0; i < 7 && !rc; ++i) rc = ialloc(S, vec[i], n3);
  if (!rc) rc = ialloc(S, &d.part, 4 * (size_t)nb);
  if (!rc) rc = ialloc(S, &d.ctrl, 1);
  if (rc) {
    implicit_release(S);
    return rc;
  }
  d.slice_off = so;
  d.row_len = rl;
  d.bcol = bc;
  d.diag_slot = dsl;
  d.A = A;
  hipError_t e = hipMemcpy(so, S->slice_off_h.data(), sizeof(int) * (S->nslice + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(rl, S->row_len_h.data(), sizeof(int) * N, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(bc, S->bcol_h.data(), sizeof(int) * S->slots, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dsl, diag.data(), sizeof(int) * N, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    implicit_release(S);
    return set_error(PBS_ERR_HIP, std::string("backward-Euler setup: ") + hipGetErrorString(e));
  }
  s->implicit = S;
  *out = S;
  return PBS_OK;
}

// Timing events: 4 per CG iteration of a replay (SpMV and update dispatches), 2 around the
// assembly or the local step.
static int cg_events(ImplicitState *S) {
  while (S->ev.size() < 4 * (size_t)kGraphIters + 2) {
    hipEvent_t ev;
    HIPCHK(hipEventCreate(&ev));
    S->ev.push_back(ev);
  }
  return PBS_OK;
}

// One graph holds kGraphIters CG iterations on the matrix of `dev`.
template <typename R> int cg_graph(pbs_solver *s, const IeDev &dev, CgGraphs &G) {
  const double key[2] = {dev.tol2, (double)dev.maxit};
  if (G.graph && std::memcmp(key, G.key, sizeof(key)) == 0) return PBS_OK;
  if (G.graph) (void)hipGraphExecDestroy(G.graph);
  G.graph = nullptr;
  std::memcpy(G.key, key, sizeof(key));
  HIPCHK(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
  hipError_t err = hipSuccess;
  for (int i = 0; i < kGraphIters && err == hipSuccess; ++i) {
    IeLaunch<R>::spmv(dev, s->st);
    IeLaunch<R>::alpha(dev, s->st);
    IeLaunch<R>::update(dev, s->st);
    IeLaunch<R>::beta(dev, s->st);
    err = hipGetLastError();
  }
  hipGraph_t gr = nullptr;
  hipError_t e2 = hipStreamEndCapture(s->st, &gr);
  if (err == hipSuccess) err = e2;
  hipGraphExec_t ex = nullptr;
  if (err == hipSuccess) err = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
  if (gr) (void)hipGraphDestroy(gr);
  if (err != hipSuccess) return set_error(PBS_ERR_HIP, std::string("CG graph capture: ") + hipGetErrorString(err));
  G.graph = ex;
  return PBS_OK;
}

// Solves dev's system from the start vector in dev.x (b, dinv already set): the start kernels,
// then graph replays until the device-side test stops; with profiling on, the same iterations are
// launched eagerly instead, the SpMV and update dispatches carrying their own start and end
// events.  c receives the final control block.
template <typename R>
int cg_solve(pbs_solver *s, ImplicitState *S, IeDev &dev, CgGraphs &G, double tol, int maxit, IeCtrl &c) {
  dev.tol2 = tol * tol;
  dev.maxit = maxit;
  const bool timed = s->prof_period > 0;
  int rc = timed ? cg_events(S) : cg_graph<R>(s, dev, G);
  if (rc) return rc;
  IeLaunch<R>::start(dev, s->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(&c, dev.ctrl, sizeof(IeCtrl), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  long replays = 0;
  const long max_replays = (maxit + kGraphIters - 1) / kGraphIters + 1;
  while (!c.done) {
    if (replays++ >= max_replays) return set_error(PBS_ERR_STATE, "CG loop did not stop within its iteration cap");
    const int k_before = c.k;
    if (timed) {
      for (int i = 0; i < kGraphIters; ++i) {
        IeLaunch<R>::spmv(dev, s->st, S->ev[4 * i], S->ev[4 * i + 1]);
        IeLaunch<R>::alpha(dev, s->st);
        IeLaunch<R>::update(dev, s->st, S->ev[4 * i + 2], S->ev[4 * i + 3]);
        IeLaunch<R>::beta(dev, s->st);
      }
      HIPCHK(hipGetLastError());
    } else {
      HIPCHK(hipGraphLaunch(G.graph, s->st));
    }
    HIPCHK(hipMemcpyAsync(&c, dev.ctrl, sizeof(IeCtrl), hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    if (timed) {  // iterations that ran: k advanced by one per completed iteration (+1 if stopped in alpha)
      const int ran = std::min(kGraphIters, c.k - k_before + (c.indefinite ? 1 : 0));
      for (int i = 0; i < ran; ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, S->ev[4 * i], S->ev[4 * i + 1]) == hipSuccess) {
          s->kt_ms[K_IE_SPMV] += ms;
          s->kt_n[K_IE_SPMV] += 1;
        }
        if (c.indefinite && i == ran - 1) break;
        if (hipEventElapsedTime(&ms, S->ev[4 * i + 2], S->ev[4 * i + 3]) == hipSuccess) {
          s->kt_ms[K_IE_UPDATE] += ms;
          s->kt_n[K_IE_UPDATE] += 1;
        }
      }
    }
  }
  if (c.indefinite) return set_error(PBS_ERR_STATE, "CG: the system matrix is not positive definite at this state");
  return PBS_OK;
}

int check_iparams(const pbs_implicit_params *ip) {
  if (!ip) return set_error(PBS_ERR_ARG, "null implicit parameters");
  if (!(ip->cg_tolerance > 0) || !(ip->cg_tolerance < 1)) return set_error(PBS_ERR_ARG, "cg_tolerance must lie in (0, 1)");
  if (ip->cg_max_iterations < 1) return set_error(PBS_ERR_ARG, "cg_max_iterations must be >= 1");
  if (!(ip->damping_alpha >= 0) || !(ip->damping_beta >= 0)) return set_error(PBS_ERR_ARG, "damping must be >= 0");
  return PBS_OK;
}

int stage_ext(pbs_solver *s, const float *ext) {
  if (!ext) return PBS_OK;
  HIPCHK(hipMemcpyAsync(s->stage, ext, sizeof(float) * 3 * (size_t)s->nn, hipMemcpyHostToDevice, s->st));
  if (s->p.precision == PBS_FP64) aos_to_planes_f(s->stage, s->m64.ext, s->m64.perm, s->nn, s->st);
  else aos_to_planes_f(s->stage, s->m32.ext, s->m32.perm, s->nn, s->st);
  HIPCHK(hipGetLastError());
  return PBS_OK;
}

template <typename R> int finish_step(pbs_solver *s, float *out) {
  HIPCHK(hipEventRecord(s->ev1, s->st));
  if (out) {
    Launch<R>::planes_to_aos(mesh<R>(s).xpos, s->stage, mesh<R>(s).perm, s->nn, s->st);
    HIPCHK(hipMemcpyAsync(out, s->stage, sizeof(float) * 3 * (size_t)s->nn, hipMemcpyDeviceToHost, s->st));
  }
  HIPCHK(hipStreamSynchronize(s->st));
  return PBS_OK;
}

template <typename R>
int implicit_step_impl(pbs_solver *s, const pbs_implicit_params *ip, const float *ext, float *out,
                       pbs_implicit_stats *stats) {
  ImplicitState *S = nullptr;
  int rc = implicit_get(s, &S);
  if (!rc) rc = stage_ext(s, ext);
  if (rc) return rc;
  DevMesh<R> &m = mesh<R>(s);
  IeDev &d = S->dev;
  IeAsm a{};
  a.h = s->p.dt;
  a.alpha = ip->damping_alpha;
  a.beta = ip->damping_beta;
  a.raw = 0;
  a.warm = ip->warm_start ? 1 : 0;
  a.has_ext = ext ? 1 : 0;
  const bool timed = s->prof_period > 0;
  if (timed) {
    rc = cg_events(S);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(s->ev0, s->st));
  hipEvent_t *ea = timed ? &S->ev[4 * kGraphIters] : nullptr;  // events around the assembly
  if (ea) HIPCHK(hipEventRecord(ea[0], s->st));
  IeLaunch<R>::assemble(m, s->td, s->dp, d, a, S->max_width, s->st);
  HIPCHK(hipGetLastError());
  if (ea) HIPCHK(hipEventRecord(ea[1], s->st));
  IeCtrl c{};
  rc = cg_solve<R>(s, S, d, S->cg_ie, ip->cg_tolerance, ip->cg_max_iterations, c);
  if (rc) return rc;
  if (ea) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ea[0], ea[1]) == hipSuccess) {
      s->kt_ms[K_IE_ASSEMBLE] += ms;
      s->kt_n[K_IE_ASSEMBLE] += 1;
    }
  }
  IeLaunch<R>::finish(m, s->dp, s->p.dt, d, s->st);
  HIPCHK(hipGetLastError());
  rc = finish_step<R>(s, out);
  if (rc) return rc;
  if (stats) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, s->ev0, s->ev1);
    std::memset(stats, 0, sizeof(*stats));
    stats->cg_iterations = c.k;
    stats->converged = c.converged;
    stats->residual0 = std::sqrt(c.rz0);
    stats->residual = std::sqrt(c.rz);
    stats->rhs_norm = std::sqrt(c.bb);
    stats->gpu_ms = ms;
  }
  return PBS_OK;
}

// ---- projective dynamics ----
template <typename R> int pd_get(pbs_solver *s, ImplicitState **out) {
  ImplicitState *S = nullptr;
  int rc = implicit_get(s, &S);
  if (rc) return rc;
  if (!S->pd.A) {
    char *A = nullptr;
    rc = ialloc(S, &A, sizeof(R) * (size_t)S->slots);
    if (!rc) rc = ialloc(S, &S->pd_s, 3 * (size_t)s->nn);
    if (!rc) rc = ialloc(S, &S->pd_inc_pos, 4 * (size_t)s->nt);
    if (rc) return rc;
    std::vector<int> pos(4 * (size_t)s->nt);
    const TilePlan &P = s->plan;
    for (size_t e = 0; e < P.inc.size(); ++e) pos[P.inc[e]] = (int)e;
    HIPCHK(hipMemcpy(S->pd_inc_pos, pos.data(), sizeof(int) * pos.size(), hipMemcpyHostToDevice));
    double *dinv = nullptr;
    rc = ialloc(S, &dinv, 3 * (size_t)s->nn);
    if (rc) return rc;
    S->pd = S->dev;  // same layout, vectors and control block; its own matrix and Jacobi diagonal
    S->pd.A = A;
    S->pd.dinv = dinv;
    S->pd.scalar = 1;
  }
  *out = S;
  return PBS_OK;
}

int check_pdparams(const pbs_pd_params *pp) {
  if (!pp) return set_error(PBS_ERR_ARG, "null projective-dynamics parameters");
  if (!(pp->stiffness > 0)) return set_error(PBS_ERR_ARG, "stiffness must be > 0");
  if (pp->iterations < 1) return set_error(PBS_ERR_ARG, "iterations must be >= 1");
  if (!(pp->cg_tolerance > 0) || !(pp->cg_tolerance < 1)) return set_error(PBS_ERR_ARG, "cg_tolerance must lie in (0, 1)");
  if (pp->cg_max_iterations < 1) return set_error(PBS_ERR_ARG, "cg_max_iterations must be >= 1");
  return PBS_OK;
}

template <typename R>
int pd_step_impl(pbs_solver *s, const pbs_pd_params *pp, const float *ext, float *out, pbs_pd_stats *stats) {
  ImplicitState *S = nullptr;
  int rc = pd_get<R>(s, &S);
  if (!rc) rc = stage_ext(s, ext);
  if (rc) return rc;
  DevMesh<R> &m = mesh<R>(s);
  IeDev &d = S->pd;
  HIPCHK(hipEventRecord(s->ev0, s->st));
  if (!S->pd_built || S->pd_stiffness != pp->stiffness || S->pd_h != (double)s->dp.h) {
    IeLaunch<R>::pd_matrix(m, s->dp, pp->stiffness, d, s->st);
    S->pd_built = true;
    S->pd_stiffness = pp->stiffness;
    S->pd_h = (double)s->dp.h;
  }
  IeLaunch<R>::pd_begin(m, s->dp, d, S->pd_s, ext ? 1 : 0, s->st);
  HIPCHK(hipGetLastError());
  long its = 0;
  int conv = 1;
  const bool timed = s->prof_period > 0;
  if (timed) {
    rc = cg_events(S);
    if (rc) return rc;
  }
  hipEvent_t *ea = timed ? &S->ev[4 * kGraphIters] : nullptr;  // events around the local step
  for (int it = 0; it < pp->iterations; ++it) {
    // local step at the current iterate (dev.x), global solve warm-started from it
    if (ea) HIPCHK(hipEventRecord(ea[0], s->st));
    IeLaunch<R>::pd_local(m, s->dp, pp->stiffness, S->pd_inc_pos, d.x, S->pd_s, d.b, s->st);
    HIPCHK(hipGetLastError());
    if (ea) HIPCHK(hipEventRecord(ea[1], s->st));
    IeCtrl c{};
    rc = cg_solve<R>(s, S, d, S->cg_pd, pp->cg_tolerance, pp->cg_max_iterations, c);
    if (rc) return rc;
    if (ea) {  // cg_solve synchronised the stream
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ea[0], ea[1]) == hipSuccess) {
        s->kt_ms[K_PD_LOCAL] += ms;
        s->kt_n[K_PD_LOCAL] += 1;
      }
    }
    its += c.k;
    conv &= c.converged;
  }
  IeLaunch<R>::pd_finish(m, s->dp, pp->damping, d, s->st);
  HIPCHK(hipGetLastError());
  rc = finish_step<R>(s, out);
  if (rc) return rc;
  if (stats) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, s->ev0, s->ev1);
    std::memset(stats, 0, sizeof(*stats));
    stats->cg_iterations = (int)its;
    stats->converged = conv;
    stats->gpu_ms = ms;
  }
  return PBS_OK;
}

template <typename R> int pd_project_impl(pbs_solver *s, float stiffness, const float *pos, double *proj) {
  ImplicitState *S = nullptr;
  int rc = pd_get<R>(s, &S);
  if (rc) return rc;
  const size_t n3 = 3 * (size_t)s->nn;
  // positions (caller order, FP32) -> solver order doubles in the iterate vector
  std::vector<double> h(n3), o(n3);
  const std::vector<int> &perm = s->plan.perm;
  for (int i = 0; i < s->nn; ++i)
    for (int k = 0; k < 3; ++k) h[3 * (size_t)i + k] = (double)pos[3 * (size_t)perm[i] + k];
  HIPCHK(hipMemcpyAsync(S->pd.x, h.data(), sizeof(double) * n3, hipMemcpyHostToDevice, s->st));
  IeLaunch<R>::pd_local(mesh<R>(s), s->dp, stiffness, S->pd_inc_pos, S->pd.x, nullptr, S->pd.b, s->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(h.data(), S->pd.b, sizeof(double) * n3, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  for (int i = 0; i < s->nn; ++i)
    for (int k = 0; k < 3; ++k) proj[3 * (size_t)perm[i] + k] = h[3 * (size_t)i + k];
  return PBS_OK;
}

template <typename R> int stiffness_impl(pbs_solver *s, int64_t nnz, double *csr_val) {
  ImplicitState *S = nullptr;
  int rc = implicit_get(s, &S);
  if (rc) return rc;
  const int N = s->nn;
  int64_t total = 0;
  for (int n = 0; n < N; ++n) total += 9L * S->row_len_h[n];
  if (nnz != total) return set_error(PBS_ERR_ARG, "nnz does not match the stiffness pattern (" + std::to_string(total) + ")");
  IeAsm a{};
  a.raw = 1;
  IeLaunch<R>::assemble(mesh<R>(s), s->td, s->dp, S->dev, a, S->max_width, s->st);
  HIPCHK(hipGetLastError());
  std::vector<R> A(9 * (size_t)S->slots);
  HIPCHK(hipMemcpyAsync(A.data(), S->dev.A, sizeof(R) * A.size(), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  // reference CSR order: row 3a+d of caller node a, columns 3b+e with b ascending in caller ids
  const std::vector<int> &perm = s->plan.perm, &iperm = s->plan.iperm;
  std::vector<int64_t> row_start(N + 1, 0);
  for (int a2 = 0; a2 < N; ++a2) row_start[a2 + 1] = row_start[a2] + 9L * S->row_len_h[iperm[a2]];
  std::vector<std::pair<int, long>> nb;
  for (int a2 = 0; a2 < N; ++a2) {
    const int n = iperm[a2];
    const long g0 = S->slice_off_h[n >> 6];
    const int lane = n & 63, len = S->row_len_h[n];
    nb.clear();
    for (int j = 0; j < len; ++j) nb.push_back({perm[S->bcol_h[(g0 + j) * 64 + lane]], g0 + j});
    std::sort(nb.begin(), nb.end());
    const int64_t rs = row_start[a2];
    const int64_t rl = 3L * len;  // values per CSR row
    for (int j = 0; j < len; ++j)
      for (int dd = 0; dd < 3; ++dd)
        for (int ee = 0; ee < 3; ++ee)
          csr_val[rs + dd * rl + 3L * j + ee] = (double)A[((size_t)nb[j].second * 9 + 3 * dd + ee) * 64 + lane];
  }
  return PBS_OK;
}

// Per-tet stiffness of the listed tets (caller ids, already checked and turned into solver ids
// in sid) through the device buffers dids (num ints) and dK (144 num floats).
template <typename R>
int tet_stiffness_run(pbs_solver *s, const float *pos, int num, const int *sid, int *dids, float *dK, float *K) {
  DevMesh<R> &m = mesh<R>(s);
  hipError_t e = hipMemcpyAsync(s->stage, pos, sizeof(float) * 3 * (size_t)s->nn, hipMemcpyHostToDevice, s->st);
  if (e == hipSuccess) {
    Launch<R>::aos_to_planes(s->stage, m.ring.b[0], m.perm, s->nn, s->st);
    e = hipMemcpyAsync(dids, sid, sizeof(int) * num, hipMemcpyHostToDevice, s->st);
  }
  if (e == hipSuccess) {
    IeLaunch<R>::tet_stiffness(m, s->dp, s->td.inst_mu, s->td.inst_lam, m.ring.b[0], num, dids, dK, s->st);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(K, dK, sizeof(float) * 144 * (size_t)num, hipMemcpyDeviceToHost, s->st);
  if (e == hipSuccess) e = hipStreamSynchronize(s->st);
  if (e != hipSuccess) return set_error(PBS_ERR_HIP, std::string("pbs_tet_stiffness: ") + hipGetErrorString(e));
  return PBS_OK;
}

int tet_ids_to_solver(pbs_solver *s, int num, const int *ids, int *sid) {
  for (int k = 0; k < num; ++k) {
    if (ids[k] < 0 || ids[k] >= s->nt) return set_error(PBS_ERR_ARG, "pbs_tet_stiffness: tet id out of range");
    sid[k] = s->plan.itperm[ids[k]];
  }
  return PBS_OK;
}

template <typename R> int tet_stiffness_impl(pbs_solver *s, const float *pos, int num, const int *ids, float *K) {
  std::vector<int> sid(std::max(num, 1));
  int rc = tet_ids_to_solver(s, num, ids, sid.data());
  if (rc) return rc;
  int *dids = nullptr;
  float *dK = nullptr;
  HIPCHK(hipMalloc(&dids, sizeof(int) * sid.size()));
  hipError_t e = hipMalloc(&dK, sizeof(float) * 144 * (size_t)std::max(num, 1));
  if (e != hipSuccess) {
    (void)hipFree(dids);
    return set_error(PBS_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  rc = tet_stiffness_run<R>(s, pos, num, sid.data(), dids, dK, K);
  (void)hipFree(dids);
  (void)hipFree(dK);
  return rc;
}

// ---- the variant of pbs_tet_stiffness that allocates nothing ----
// Its device buffers (the listed tets' solver ids, then their 12 x 12 matrices) live in one
// file-scope device array, one copy per GPU, held until the process ends.  Host code never uses
// the array's name as an address: the address is looked up per device with
// hipGetSymbolAddress on that device's first call and kept in scratch_ptr.
constexpr int kScratchTets = 8192;  // tets per call
constexpr size_t kScratchIdBytes = (sizeof(int) * kScratchTets + 255) / 256 * 256;
constexpr size_t kScratchBytes = kScratchIdBytes + sizeof(float) * 144 * (size_t)kScratchTets;  // 4.75 MB
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void *scratch_ptr[kMaxDevices] = {};

int scratch_address(int device, unsigned char **out) {
  if (device < 0 || device >= kMaxDevices) return set_error(PBS_ERR_STATE, "device index out of range");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[device]) {
    int cur = -1;
    HIPCHK(hipGetDevice(&cur));
    if (cur != device) HIPCHK(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return set_error(PBS_ERR_HIP, std::string("scratch address (hipGetSymbolAddress): ") + hipGetErrorString(e));
    scratch_ptr[device] = p;
  }
  *out = (unsigned char *)scratch_ptr[device];
  return PBS_OK;
}

template <typename R> int tet_stiffness_noalloc_impl(pbs_solver *s, const float *pos, int num, const int *ids, float *K) {
  int sid[kScratchTets];
  int rc = tet_ids_to_solver(s, num, ids, sid);
  if (rc) return rc;
  unsigned char *base = nullptr;
  rc = scratch_address(s->device, &base);
  if (rc) return rc;
  return tet_stiffness_run<R>(s, pos, num, sid, (int *)base, (float *)(base + kScratchIdBytes), K);
}

}  // namespace

extern "C" int pbs_implicit_default_params(pbs_implicit_params *p) {
  if (!p) return set_error(PBS_ERR_ARG, "null argument");
  std::memset(p, 0, sizeof(*p));
  p->damping_alpha = 0.02f;  // m_dumpingAlpha, m_dumpingBelta (CUDAInterface.cu:644)
  p->damping_beta = 0.02f;
  p->cg_tolerance = 1e-6;
  p->cg_max_iterations = 10000;  // conjugateGradient's cap (CUDALinearSolvers.cu)
  p->warm_start = 1;
  return PBS_OK;
}

extern "C" int pbs_implicit_step(pbs_solver *s, const pbs_implicit_params *ip, const float *ext, float *out,
                                 pbs_implicit_stats *stats) {
  if (!s) return set_error(PBS_ERR_ARG, "null handle");
  int rc = check_iparams(ip);
  if (rc) return rc;
  return s->p.precision == PBS_FP64 ? implicit_step_impl<double>(s, ip, ext, out, stats)
                                    : implicit_step_impl<float>(s, ip, ext, out, stats);
}

extern "C" int pbs_tet_stiffness(pbs_solver *s, const float *pos, int num, const int *tet_ids, float *K) {
  if (!s || !pos || num < 0 || (num > 0 && (!tet_ids || !K))) return set_error(PBS_ERR_ARG, "pbs_tet_stiffness: bad arguments");
  if (s->ninst != 1) return set_error(PBS_ERR_STATE, "pbs_tet_stiffness: single-mesh handles only");
  if (num == 0) return PBS_OK;
  return s->p.precision == PBS_FP64 ? tet_stiffness_impl<double>(s, pos, num, tet_ids, K)
                                    : tet_stiffness_impl<float>(s, pos, num, tet_ids, K);
}

extern "C" int pbs_tet_stiffness_noalloc(pbs_solver *s, const float *pos, int num, const int *tet_ids, float *K) {
  if (!s || !pos || num < 0 || (num > 0 && (!tet_ids || !K)))
    return set_error(PBS_ERR_ARG, "pbs_tet_stiffness_noalloc: bad arguments");
  if (num > kScratchTets)
    return set_error(PBS_ERR_ARG, "pbs_tet_stiffness_noalloc: at most " + std::to_string(kScratchTets) +
                                      " tets per call (" + std::to_string(num) + " given)");
  if (s->ninst != 1) return set_error(PBS_ERR_STATE, "pbs_tet_stiffness_noalloc: single-mesh handles only");
  if (num == 0) return PBS_OK;
  return s->p.precision == PBS_FP64 ? tet_stiffness_noalloc_impl<double>(s, pos, num, tet_ids, K)
                                    : tet_stiffness_noalloc_impl<float>(s, pos, num, tet_ids, K);
}

extern "C" int pbs_stiffness_matrix(pbs_solver *s, int64_t nnz, double *csr_val) {
  if (!s || !csr_val) return set_error(PBS_ERR_ARG, "null argument");
  return s->p.precision == PBS_FP64 ? stiffness_impl<double>(s, nnz, csr_val) : stiffness_impl<float>(s, nnz, csr_val);
}

extern "C" int pbs_pd_default_params(pbs_pd_params *p) {
  if (!p) return set_error(PBS_ERR_ARG, "null argument");
  std::memset(p, 0, sizeof(*p));
  p->stiffness = 1e6f;  // m_stiffness, m_iterations (ProjDynamic.cpp:4)
  p->iterations = 5;
  p->damping = 0.99f;   // ProjDynamic.cpp:477
  // the unknowns are positions: the fixed nodes' rows (mass 1e8) carry most of |D^-1/2 b|, so the
  // free nodes need a tighter relative tolerance than the backward-Euler velocity solve
  p->cg_tolerance = 1e-10;
  p->cg_max_iterations = 10000;
  return PBS_OK;
}

extern "C" int pbs_pd_step(pbs_solver *s, const pbs_pd_params *p, const float *ext, float *out, pbs_pd_stats *st) {
  if (!s) return set_error(PBS_ERR_ARG, "null handle");
  int rc = check_pdparams(p);
  if (rc) return rc;
  return s->p.precision == PBS_FP64 ? pd_step_impl<double>(s, p, ext, out, st) : pd_step_impl<float>(s, p, ext, out, st);
}

extern "C" int pbs_pd_project(pbs_solver *s, float stiffness, const float *positions, double *proj) {
  if (!s || !positions || !proj) return set_error(PBS_ERR_ARG, "null argument");
  if (!(stiffness > 0)) return set_error(PBS_ERR_ARG, "stiffness must be > 0");
  return s->p.precision == PBS_FP64 ? pd_project_impl<double>(s, stiffness, positions, proj)
                                    : pd_project_impl<float>(s, stiffness, positions, proj);
}
