// pbs_capi.cpp -- C ABI of the descent solver (include/phybs.h): handle lifetime, state I/O,
// the step driver and the GPU material calls.  Replaces CUDAInterface's host side
// (CUDAInterface.cu:641-1007).  Everything runs on the handle's stream; the only host round
// trip inside a step is one read of the 128-byte device control block after the planned passes
// (the reference reads an energy every 8 iterations, CUDAInterface.cu:942-944).
//
// Inside the solver nodes and tets are renumbered (Morton order, pbs_tiles.cpp); every array that
// crosses the ABI is in the caller's order and is permuted on the device or the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "pbs_host.h"
#include "pbs_internal.h"
#include "pbs_learn.h"
#include "pbs_solver.h"
#include "pbs_tiles.h"
#include "phybs.h"

namespace pbs {
static thread_local std::string g_err;
int set_error(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
}  // namespace pbs

using namespace pbs;

static int validate_params(const pbs_params *p) {
  if (!p) return set_error(PBS_ERR_ARG, "null params");
  if (p->precision != PBS_FP32 && p->precision != PBS_FP64) return set_error(PBS_ERR_ARG, "precision must be PBS_FP32 or PBS_FP64");
  if (p->semantics != PBS_SEM_REF_GPU && p->semantics != PBS_SEM_REF_CPU) return set_error(PBS_ERR_ARG, "bad semantics");
  if (p->material != PBS_MAT_NEOHOOKEAN && p->material != PBS_MAT_STRETCH) return set_error(PBS_ERR_ARG, "bad material");
  if (!(p->dt > 0) || p->hd_refresh <= 0 || p->energy_check <= 0 || p->iterations < 0 || p->profile_n < 0 || p->profile_n > 8)
    return set_error(PBS_ERR_ARG, "bad loop parameters");
  if (!(p->youngs > 0) || !(p->nu > -1.0f && p->nu < 0.5f) || !(p->density > 0)) return set_error(PBS_ERR_ARG, "bad material constants");
  if (p->relgrad_tol < 0 || (p->relgrad_tol > 0 && p->max_iterations <= 0)) return set_error(PBS_ERR_ARG, "bad rel-grad settings");
  if (p->block_size != 0 && p->block_size != 64 && p->block_size != 128 && p->block_size != 256)
    return set_error(PBS_ERR_ARG, "block_size must be 0, 64, 128 or 256");
  if (p->tile_tets != 0 && p->tile_tets != -1 && (p->tile_tets < 16 || p->tile_tets > 4096))
    return set_error(PBS_ERR_ARG, "tile_tets must be 0, -1 or in [16, 4096]");
  if (p->graph_passes < -1 || p->graph_passes > 1024) return set_error(PBS_ERR_ARG, "graph_passes must be in [-1, 1024]");
  if (!(p->mixed_switch >= 0 && p->mixed_switch < 1)) return set_error(PBS_ERR_ARG, "mixed_switch must lie in [0, 1)");
  if (p->mixed_switch > 0 && p->precision != PBS_FP64) return set_error(PBS_ERR_ARG, "mixed_switch needs PBS_FP64");
  return PBS_OK;
}

static void make_devparams(pbs_solver *s) {
  const pbs_params &p = s->p;
  DevParams &d = s->dp;
  d.nn = s->nn;
  d.nt = s->nt;
  d.h = p.dt;
  // mu = 0.5 E / (1 + nu), lambda = 2 mu / (1 - 2 nu) (CUDAInterface.cu:704-705)
  d.mu = (float)(0.5 * (double)p.youngs / (double)(1 + p.nu));
  d.lam = d.mu * 2 / (1 - 2 * p.nu);
  d.density = p.density;
  d.fixed_mass = p.fixed_mass;
  for (int i = 0; i < 3; ++i) d.grav[i] = p.grav[i];
  d.alpha0 = p.alpha0;
  d.alpha_decay = p.alpha_decay;
  d.alpha_min = p.alpha_min;
  d.hd_refresh = p.hd_refresh;
  d.energy_check = p.energy_check;
  d.damping = p.damping;
  d.init_momentum = p.init_momentum;
  d.profile_n = p.profile_n;
  for (int i = 0; i < 8; ++i) {
    d.profile_k[i] = p.profile_k[i];
    d.profile_rho[i] = p.profile_rho[i];
  }
  d.semantics = p.semantics;
  d.material = p.material;
  d.clamp = p.semantics == PBS_SEM_REF_CPU ? 2 : 1;
  d.ext_scale = p.semantics == PBS_SEM_REF_CPU ? 2.0f : 1.0f;
  d.has_ext = 0;
  d.iterations = p.relgrad_tol > 0 ? p.max_iterations : p.iterations;
  d.relgrad_tol = p.relgrad_tol;
  d.f32math = 0;
  s->bs = p.block_size ? p.block_size : 256;
  s->ntb = (s->nt + s->bs - 1) / s->bs;
  s->nnb = (s->nn + s->bs - 1) / s->bs;
}

// ---- kernel timing ----
// A pair of events for one timed launch (kernel id kid, descent-pass ordinal s->pass_ordinal).
static int timed_slot(pbs_solver *s, int kid) {
  if (!s->prof_period) return -1;
  if (s->pool_used + 2 > (int)s->pool.size()) {
    for (int i = 0; i < 64; ++i) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return -1;
      s->pool.push_back(e);
    }
  }
  int i = s->pool_used;
  s->pool_used += 2;
  s->pending.push_back({kid, i, s->pass_ordinal});
  return i;
}
// Times the stream operations between timed_begin and timed_end (several launches).
static int timed_begin(pbs_solver *s, int kid) {
  const int i = timed_slot(s, kid);
  if (i >= 0) (void)hipEventRecord(s->pool[i], s->st);
  return i;
}
static void timed_end(pbs_solver *s, int i) {
  if (i >= 0) (void)hipEventRecord(s->pool[i + 1], s->st);
}
// Adds the recorded launch times to the per-kernel totals.  Descent-pass launches with an
// ordinal >= executed ran after the loop had stopped (they return at once) and are not counted.
static void timed_collect(pbs_solver *s, long executed) {
  for (auto &pr : s->pending) {
    float ms = 0.f;
    if (pr.pass >= executed) continue;
    if (hipEventElapsedTime(&ms, s->pool[pr.ev], s->pool[pr.ev + 1]) == hipSuccess) {
      s->kt_ms[pr.kid] += ms;
      s->kt_n[pr.kid] += 1;
    }
  }
  s->pending.clear();
  s->pool_used = 0;
}

static void graph_release(pbs_solver *s) {
  for (auto &g : s->graph) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
    g.passes = 0;
  }
}

// ---- state helpers ----
template <typename R> static int upload_planes(pbs_solver *s, const float *aos, R *planes) {
  HIPCHK(hipMemcpyAsync(s->stage, aos, sizeof(float) * 3 * (size_t)s->nn, hipMemcpyHostToDevice, s->st));
  Launch<R>::aos_to_planes(s->stage, planes, mesh<R>(s).perm, s->nn, s->st);
  HIPCHK(hipGetLastError());
  return PBS_OK;
}
template <typename R> static int download_planes(pbs_solver *s, const R *planes, float *aos) {
  Launch<R>::planes_to_aos(planes, s->stage, mesh<R>(s).perm, s->nn, s->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(aos, s->stage, sizeof(float) * 3 * (size_t)s->nn, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return PBS_OK;
}

template <typename R> static int reset_impl(pbs_solver *s) {
  DevMesh<R> &m = mesh<R>(s);
  int rc = upload_planes<R>(s, s->rest_h.data(), m.xpos);
  if (rc) return rc;
  HIPCHK(hipMemsetAsync(m.v, 0, sizeof(R) * 3 * (size_t)s->nn, s->st));
  HIPCHK(hipMemsetAsync(m.vl, 0, sizeof(R) * 3 * (size_t)s->nn, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return PBS_OK;
}

// Per-instance Lame constants, with the reference's formulas (CUDAInterface.cu:704-705) in the
// same arithmetic as the single-mesh constants of make_devparams.
static int upload_lame(pbs_solver *s) {
  std::vector<float> f(2 * (size_t)s->ninst);
  for (int i = 0; i < s->ninst; ++i) {
    const float E = s->inst_youngs[i], nu = s->inst_nu[i];
    const float mu = (float)(0.5 * (double)E / (double)(1 + nu));
    f[i] = mu;
    f[s->ninst + i] = mu * 2 / (1 - 2 * nu);
  }
  HIPCHK(hipMemcpyAsync((void *)s->td.inst_mu, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice, s->st));
  return PBS_OK;
}

template <typename R> static int alloc_impl(pbs_solver *s) {
  DevMesh<R> &m = mesh<R>(s);
  const size_t N = s->nn, T = std::max(s->nt, 1);
  int rc = 0;
  rc |= dalloc(s, &m.tets, T);
  rc |= dalloc(s, &m.rest, 3 * N);
  rc |= dalloc(s, &m.dmi, 9 * T);
  rc |= dalloc(s, &m.vol, T);
  rc |= dalloc(s, &m.inc_off, N + 1);
  rc |= dalloc(s, &m.inc, 4 * T);
  rc |= dalloc(s, &m.perm, N);
  rc |= dalloc(s, &m.mask, N);
  rc |= dalloc(s, &m.mass, N);
  rc |= dalloc(s, &m.mnode, N);
  rc |= dalloc(s, &m.gforce, 3 * N);
  rc |= dalloc(s, &m.ext, 3 * N);
  rc |= dalloc(s, &m.hd, 3 * N);
  rc |= dalloc(s, &m.frec, 12 * T);
  rc |= dalloc(s, &m.hrec, 12 * T);
  rc |= dalloc(s, &m.xt, 3 * N);
  rc |= dalloc(s, &m.xpos, 3 * N);
  rc |= dalloc(s, &m.v, 3 * N);
  rc |= dalloc(s, &m.vl, 3 * N);
  for (int i = 0; i < 4; ++i) rc |= dalloc(s, &m.ring.b[i], 3 * N);
  // partial-sum slots for the smallest supported block (64 threads)
  rc |= dalloc(s, &m.etet, T / 64 + 2);
  rc |= dalloc(s, &m.enode, N / 64 + 2);
  rc |= dalloc(s, &m.gpart, N / 64 + 2);
  // tile pass
  const TilePlan &P = s->plan;
  const size_t H = std::max<long>(P.halo_total, 1);
  TileInfo *tiles;
  int *halo, *bnd_off, *hslot;
  uint32_t *tlp;
  unsigned short *toff;
  rc |= dalloc(s, &tiles, (size_t)P.ntiles);
  rc |= dalloc(s, &halo, H);
  rc |= dalloc(s, &tlp, 4 * T);
  rc |= dalloc(s, &toff, P.tile_off.size());
  rc |= dalloc(s, &bnd_off, N + 1);
  rc |= dalloc(s, &hslot, H);
  rc |= dalloc(s, &s->td.tpart, (size_t)P.ntiles);
  rc |= dalloc(s, &s->td.npart, 2 * (size_t)std::max(s->td.nblk, 1) + 2);
  rc |= dalloc(s, &s->td.fpart, (size_t)std::max(s->td.nblk, 1));
  rc |= dalloc(s, &m.fown, 3 * N);
  rc |= dalloc(s, &m.hpart, 3 * H);
  rc |= dalloc(s, &m.fownh, 3 * N);
  rc |= dalloc(s, &m.hparth, 3 * H);
  s->td.tiles = tiles;
  s->td.halo = halo;
  s->td.tet_lp = (const uint4 *)tlp;
  s->td.tile_off = toff;
  s->td.bnd_off = bnd_off;
  s->td.halo_slot = hslot;
  rc |= dalloc(s, &m.ctrl, (size_t)s->ninst);
  {
    int *a;
    float *f;
    const size_t I = (size_t)s->ninst;
    rc |= dalloc(s, &a, 4 * (I + 1) + 2 * (size_t)std::max(s->td.nblk, 1));
    s->td.inst_node_off = a;
    s->td.inst_tile_off = a + (I + 1);
    s->td.inst_tet_off = a + 2 * (I + 1);
    s->td.inst_blk_off = a + 3 * (I + 1);
    s->td.blk_inst = a + 4 * (I + 1);
    s->td.blk_start = a + 4 * (I + 1) + std::max(s->td.nblk, 1);
    rc |= dalloc(s, &f, 2 * I);
    s->td.inst_mu = f;
    s->td.inst_lam = f + I;
  }
  rc |= dalloc(s, &m.flag, 1);
  rc |= dalloc(s, &m.tete, T);
  rc |= dalloc(s, &s->stage, 3 * N);
  rc |= dalloc(s, &s->stage64, 3 * N);
  return rc ? PBS_ERR_HIP : PBS_OK;
}

template <typename R> static int setup_impl(pbs_solver *s, const int *mask) {
  DevMesh<R> &m = mesh<R>(s);
  const TilePlan &P = s->plan;
  const size_t N = s->nn;
  std::vector<float> planes(3 * N);
  for (size_t i = 0; i < N; ++i)
    for (int d = 0; d < 3; ++d) planes[d * N + i] = s->rest_h[3 * (size_t)P.perm[i] + d];
  if (s->nt > 0) HIPCHK(hipMemcpyAsync(m.tets, P.tets_new.data(), sizeof(int) * 4 * (size_t)s->nt, hipMemcpyHostToDevice, s->st));
  HIPCHK(hipMemcpyAsync(m.rest, planes.data(), sizeof(float) * 3 * N, hipMemcpyHostToDevice, s->st));
  std::vector<int> mk(N, 0);
  if (mask)
    for (size_t i = 0; i < N; ++i) mk[i] = mask[P.perm[i]] ? 1 : 0;
  HIPCHK(hipMemcpyAsync(m.mask, mk.data(), sizeof(int) * N, hipMemcpyHostToDevice, s->st));
  HIPCHK(hipMemcpyAsync(m.perm, P.perm.data(), sizeof(int) * N, hipMemcpyHostToDevice, s->st));
  HIPCHK(hipMemcpyAsync(m.inc_off, P.inc_off.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice, s->st));
  if (s->nt > 0) HIPCHK(hipMemcpyAsync(m.inc, P.inc.data(), sizeof(int) * 4 * (size_t)s->nt, hipMemcpyHostToDevice, s->st));
  HIPCHK(hipMemcpyAsync((void *)s->td.tiles, P.tiles.data(), sizeof(TileInfo) * P.tiles.size(), hipMemcpyHostToDevice, s->st));
  {
    const size_t I = (size_t)s->ninst;
    std::vector<int> a(4 * (I + 1) + 2 * (size_t)std::max(s->td.nblk, 1), 0);
    for (size_t i = 0; i <= I; ++i) {
      a[i] = P.inst_node_off[i];
      a[(I + 1) + i] = P.inst_tile_off[i];
      a[2 * (I + 1) + i] = P.inst_tet_off[i];
    }
    int b = 0;
    for (size_t i = 0; i < I; ++i) {
      a[3 * (I + 1) + i] = b;
      for (int n = P.inst_node_off[i]; n < P.inst_node_off[i + 1]; n += 256, ++b) {
        a[4 * (I + 1) + b] = (int)i;
        a[4 * (I + 1) + std::max(s->td.nblk, 1) + b] = n;
      }
    }
    a[3 * (I + 1) + I] = b;
    HIPCHK(hipMemcpyAsync((void *)s->td.inst_node_off, a.data(), sizeof(int) * a.size(), hipMemcpyHostToDevice, s->st));
    int rc = upload_lame(s);
    if (rc) return rc;
  }
  if (!P.halo.empty())
    HIPCHK(hipMemcpyAsync((void *)s->td.halo, P.halo.data(), sizeof(int) * P.halo.size(), hipMemcpyHostToDevice, s->st));
  if (s->nt > 0)
    HIPCHK(hipMemcpyAsync((void *)s->td.tet_lp, P.tet_lp.data(), sizeof(uint32_t) * P.tet_lp.size(), hipMemcpyHostToDevice, s->st));
  HIPCHK(hipMemcpyAsync((void *)s->td.tile_off, P.tile_off.data(), sizeof(uint16_t) * P.tile_off.size(), hipMemcpyHostToDevice, s->st));
  HIPCHK(hipMemcpyAsync((void *)s->td.bnd_off, P.bnd_off.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice, s->st));
  if (!P.halo_slot.empty())
    HIPCHK(hipMemcpyAsync((void *)s->td.halo_slot, P.halo_slot.data(), sizeof(int) * P.halo_slot.size(), hipMemcpyHostToDevice, s->st));
  HIPCHK(hipMemsetAsync(m.fownh, 0, sizeof(float) * 3 * N, s->st));
  HIPCHK(hipMemsetAsync(m.flag, 0, sizeof(int), s->st));
  HIPCHK(hipMemsetAsync(m.hd, 0, sizeof(float) * 3 * N, s->st));
  HIPCHK(hipMemsetAsync(m.ext, 0, sizeof(float) * 3 * N, s->st));
  Launch<R>::setup_tets(m, s->dp, 256, s->st);
  HIPCHK(hipGetLastError());
  Launch<R>::setup_nodes(m, s->dp, 256, s->st);
  HIPCHK(hipGetLastError());
  int flag = 0;
  HIPCHK(hipMemcpyAsync(&flag, m.flag, sizeof(int), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  if (flag & 1) return set_error(PBS_ERR_MESH, "tet index out of range");
  if (flag & 2) return set_error(PBS_ERR_MESH, "rest tet with det(Dm) <= 0 (inverted or degenerate)");
  return reset_impl<R>(s);
}

static void free_all(pbs_solver *s) {
  if (s->st) (void)hipStreamSynchronize(s->st);
  for (void *q : s->allocs) (void)hipFree(q);
  s->allocs.clear();
  for (auto e : s->pool) (void)hipEventDestroy(e);
  s->pool.clear();
  graph_release(s);
  learn_release(s->learn);
  s->learn = nullptr;
  implicit_release(s->implicit);
  s->implicit = nullptr;
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->st) (void)hipStreamDestroy(s->st);
  s->st = nullptr;
}

extern "C" int pbs_abi_version(void) { return PBS_ABI_VERSION; }
extern "C" const char *pbs_last_error(void) { return g_err.c_str(); }

extern "C" int pbs_default_params(pbs_params *p) {
  if (!p) return set_error(PBS_ERR_ARG, "null params");
  std::memset(p, 0, sizeof(*p));
  p->dt = 0.03f;
  p->grav[0] = 0.0f;
  p->grav[1] = -9.8f;
  p->grav[2] = 0.0f;
  p->youngs = 1e5f;
  p->nu = 0.45f;
  p->density = 1000.0f;
  p->fixed_mass = 1e8f;
  p->alpha0 = 0.1f;
  p->alpha_decay = 0.7f;
  p->alpha_min = 0.01f;
  p->hd_refresh = 32;
  p->energy_check = 8;
  p->damping = 0.99f;
  p->init_momentum = 0.2f;
  p->profile_n = 3;
  p->profile_k[0] = 1;
  p->profile_k[1] = 7;
  p->profile_k[2] = 10;
  p->profile_rho[0] = 0.96f;
  p->profile_rho[1] = 0.991f;
  p->profile_rho[2] = 0.999f;
  p->semantics = PBS_SEM_REF_GPU;
  p->precision = PBS_FP32;
  p->material = PBS_MAT_NEOHOOKEAN;
  p->iterations = 96;
  p->relgrad_tol = 0.0;
  p->max_iterations = 20000;
  p->block_size = 0;
  return PBS_OK;
}

template <typename R> static int set_positions_impl(pbs_solver *s, const float *xyz) {
  int rc = upload_planes<R>(s, xyz, mesh<R>(s).xpos);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s->st));
  return PBS_OK;
}

// Creates a solver for ninst independent meshes (1 for pbs_create): instance i owns nodes
// [node_off[i], node_off[i+1]) and tets [tet_off[i], tet_off[i+1]), tets hold global node ids.
// tile_tets: one tile size for all instances (0: chosen from the mesh size); inst_tiles, when
// not null, one per instance instead.
static int create_impl(int ninst, const int *node_off, const int *tet_off, const float *nodes, const float *rest,
                       const int *mask, const int *tets, const float *youngs, const float *nu, const pbs_params *p,
                       int tile_tets, const int *inst_tiles, pbs_solver **out) {
  const int nn = node_off[ninst], nt = tet_off[ninst];
  pbs_solver *s = new pbs_solver();
  s->nn = nn;
  s->nt = nt;
  s->ninst = ninst;
  s->inst_youngs.assign(youngs, youngs + ninst);
  s->inst_nu.assign(nu, nu + ninst);
  s->p = *p;
  make_devparams(s);
  s->rest_h.assign(rest, rest + 3 * (size_t)nn);
  auto fail = [&](int code) {
    free_all(s);
    delete s;
    return code;
  };
  if (hipGetDevice(&s->device) != hipSuccess) return fail(set_error(PBS_ERR_HIP, "no HIP device"));
  {
    int maxlds = 65536;
    if (hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device) != hipSuccess) maxlds = 65536;
    s->lds_budget = (size_t)maxlds;
  }
  // Tile size (tets per tile): 256, halved while the mesh would give fewer than 1024 tiles, and
  // while the tile pass could need more than 40 KiB of LDS (four workgroups per 160 KiB CU); a
  // plan with tile_tets = t has at most t tets and max(64, t / 2) local nodes per tile.  FP64
  // neo-Hookean arithmetic (not the mixed mode) takes 224 instead of 256: 27.0 KB of LDS per
  // workgroup instead of 30.8 KB lets 6 workgroups share a CU, as many as the kernel's 6 waves
  // per SIMD allow (FP64 tile pass 135.8 -> 131.3 us at 6.3 M tets; 224-tet tiles alone, at 5
  // workgroups per CU, take 137.7 us).
  const bool f64 = p->precision == PBS_FP64;
  auto lds_need = [&](int lcap, int tcap) {
    TileDev t{};
    t.lcap = std::max(lcap, 1);
    t.tcap = std::max(tcap, 1);
    return f64 ? Launch<double>::tile_lds(t) : Launch<float>::tile_lds(t);
  };
  int tt = tile_tets;
  if (tt == 0) {
    tt = 256;
    while (tt > 64 && (nt / tt < 1024 || lds_need(std::max(64, tt / 2), tt) > std::min<size_t>(s->lds_budget, 40960))) tt /= 2;
    if (tt == 256 && f64 && p->material == 0 && !(p->mixed_switch > 0)) tt = 224;
  }
  std::string msg;
  const bool planned = inst_tiles ? build_batch_plan(ninst, node_off, tet_off, rest, tets, inst_tiles, s->plan, msg)
                                  : build_batch_plan(ninst, node_off, tet_off, rest, tets, tt, s->plan, msg);
  if (!planned) return fail(set_error(PBS_ERR_MESH, msg));
  if (inst_tiles) tt = s->plan.tile_tets;
  if (lds_need(s->plan.lcap, s->plan.tcap) > s->lds_budget)
    return fail(set_error(PBS_ERR_ARG, "tile too large for the LDS of one workgroup: use a smaller tile_tets"));
  s->p.tile_tets = tt;
  s->td.ntiles = s->plan.ntiles;
  s->td.lcap = std::max(s->plan.lcap, 1);
  s->td.tcap = std::max(s->plan.tcap, 1);
  // one thread per tet of a full tile, in whole wavefronts, at most 256
  s->td.threads = std::min(256, std::max(64, (s->plan.tile_tets + 63) / 64 * 64));
  s->td.halo_total = s->plan.halo_total;
  s->td.ninst = ninst;
  s->td.nblk = 0;
  for (int i = 0; i < ninst; ++i) s->td.nblk += (node_off[i + 1] - node_off[i] + 255) / 256;
  s->ctrl_h.assign(ninst, Ctrl{});
  if (hipStreamCreate(&s->st) != hipSuccess) return fail(set_error(PBS_ERR_HIP, "hipStreamCreate failed"));
  if (hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess)
    return fail(set_error(PBS_ERR_HIP, "hipEventCreate failed"));
  int rc = p->precision == PBS_FP64 ? alloc_impl<double>(s) : alloc_impl<float>(s);
  if (rc) return fail(rc);
  rc = p->precision == PBS_FP64 ? setup_impl<double>(s, mask) : setup_impl<float>(s, mask);
  if (rc) return fail(rc);
  if (nodes) {
    rc = p->precision == PBS_FP64 ? set_positions_impl<double>(s, nodes) : set_positions_impl<float>(s, nodes);
    if (rc) return fail(rc);
  }
  *out = s;
  return PBS_OK;
}

// Time of one descent pass (microseconds) of a solver built with the given tile sizes: 8
// untimed passes from the initial state, then 32 timed ones.  Negative when the solver cannot be
// built with them (e.g. a tile exceeds the LDS).
static double time_tiles(int ninst, const int *node_off, const int *tet_off, const float *nodes, const float *rest,
                         const int *mask, const int *tets, const float *youngs, const float *nu, const pbs_params *p,
                         int tt, const int *inst_tiles) {
  pbs_solver *s = nullptr;
  pbs_params q = *p;
  q.iterations = 8;
  q.relgrad_tol = 0.0;
  if (create_impl(ninst, node_off, tet_off, nodes, rest, mask, tets, youngs, nu, &q, tt, inst_tiles, &s)) return -1.0;
  pbs_step_stats st{};
  int rc = pbs_step(s, nullptr, nullptr, &st);
  q.iterations = 32;
  if (!rc) rc = pbs_set_params(s, &q);
  if (!rc) rc = pbs_reset(s);
  if (!rc && nodes) rc = pbs_set_positions(s, nodes);
  if (!rc) rc = pbs_step(s, nullptr, nullptr, &st);
  pbs_destroy(s);
  if (rc || st.passes <= 0) return -1.0;
  return 1e3 * st.gpu_ms / st.passes;
}

// Autotuner: times each candidate tile size on the whole mesh (or batch) and keeps the fastest.
// A batch of at least 8 instances is also tuned per instance: the instances are split by tet
// count into 4 classes of equal count, each class alone (a batch of its instances) is timed at
// every candidate size, and the batch with each instance at its class's fastest size is timed
// against the fastest uniform size; the faster of the two is built.
static int create_tuned(int ninst, const int *node_off, const int *tet_off, const float *nodes, const float *rest,
                        const int *mask, const int *tets, const float *youngs, const float *nu, const pbs_params *p,
                        pbs_solver **out) {
  static const int cand[] = {64, 128, 256, 512};
  std::vector<int> tiles;
  std::vector<double> us;
  int best = -1;
  double best_us = 0.0;
  for (int tt : cand) {
    const double t = time_tiles(ninst, node_off, tet_off, nodes, rest, mask, tets, youngs, nu, p, tt, nullptr);
    if (t < 0) continue;
    tiles.push_back(tt);
    us.push_back(t);
    if (best < 0 || t < best_us) {
      best = tt;
      best_us = t;
    }
  }
  if (best < 0) return set_error(PBS_ERR_ARG, "autotuner: no tile size fits this mesh");
  std::vector<int> inst_tile;
  std::vector<int> class_tile;
  if (ninst >= 8) {
    constexpr int K = 4;
    std::vector<int> order(ninst);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      return tet_off[a + 1] - tet_off[a] < tet_off[b + 1] - tet_off[b];
    });
    inst_tile.assign(ninst, best);
    for (int c = 0; c < K; ++c) {
      std::vector<int> mem(order.begin() + (size_t)c * ninst / K, order.begin() + (size_t)(c + 1) * ninst / K);
      std::sort(mem.begin(), mem.end());
      // the class as a batch of its own: instance arrays copied, node ids re-based
      std::vector<int> no(1, 0), to(1, 0), msk, tt4;
      std::vector<float> nd, rs, E, V;
      for (int i : mem) {
        const int n0 = node_off[i], n1 = node_off[i + 1], t0 = tet_off[i], t1 = tet_off[i + 1];
        for (long k = 4L * t0; k < 4L * t1; ++k) tt4.push_back(tets[k] - n0 + no.back());
        rs.insert(rs.end(), rest + 3L * n0, rest + 3L * n1);
        if (nodes) nd.insert(nd.end(), nodes + 3L * n0, nodes + 3L * n1);
        if (mask) msk.insert(msk.end(), mask + n0, mask + n1);
        E.push_back(youngs[i]);
        V.push_back(nu[i]);
        no.push_back(no.back() + n1 - n0);
        to.push_back(to.back() + t1 - t0);
      }
      int cbest = best;
      double cbest_us = -1.0;
      for (int tt : cand) {
        const double t = time_tiles((int)mem.size(), no.data(), to.data(), nodes ? nd.data() : nullptr, rs.data(),
                                    mask ? msk.data() : nullptr, tt4.data(), E.data(), V.data(), p, tt, nullptr);
        if (t >= 0 && (cbest_us < 0 || t < cbest_us)) {
          cbest = tt;
          cbest_us = t;
        }
      }
      class_tile.push_back(cbest);
      for (int i : mem) inst_tile[i] = cbest;
    }
    if (std::all_of(class_tile.begin(), class_tile.end(), [&](int t) { return t == class_tile[0]; })) {
      if (class_tile[0] == best) inst_tile.clear();  // the same as the fastest uniform size
    }
    if (!inst_tile.empty()) {
      const double t = time_tiles(ninst, node_off, tet_off, nodes, rest, mask, tets, youngs, nu, p, 0,
                                  inst_tile.data());
      tiles.push_back(-2);
      us.push_back(t);
      if (!(t >= 0 && t < best_us)) inst_tile.clear();
    }
  }
  int rc = create_impl(ninst, node_off, tet_off, nodes, rest, mask, tets, youngs, nu, p, best,
                       inst_tile.empty() ? nullptr : inst_tile.data(), out);
  if (rc) return rc;
  (*out)->tune_tiles = tiles;
  (*out)->tune_us = us;
  (*out)->tune_class_tiles = class_tile;
  return PBS_OK;
}

static int create_any(int ninst, const int *node_off, const int *tet_off, const float *nodes, const float *rest,
                      const int *mask, const int *tets, const float *youngs, const float *nu, const pbs_params *p,
                      pbs_solver **out) {
  if (p->tile_tets < 0) return create_tuned(ninst, node_off, tet_off, nodes, rest, mask, tets, youngs, nu, p, out);
  return create_impl(ninst, node_off, tet_off, nodes, rest, mask, tets, youngs, nu, p, p->tile_tets, nullptr, out);
}

extern "C" int pbs_create(int nn, const float *nodes, const float *rest, const int *mask, int nt, const int *tets,
                          const pbs_params *p, pbs_solver **out) {
  if (!out) return set_error(PBS_ERR_ARG, "null out");
  *out = nullptr;
  if (nn <= 0 || nt < 0 || !rest || (nt > 0 && !tets)) return set_error(PBS_ERR_ARG, "pbs_create: bad mesh arguments");
  int rc = validate_params(p);
  if (rc) return rc;
  for (long i = 0; i < 4L * nt; ++i)
    if (tets[i] < 0 || tets[i] >= nn) return set_error(PBS_ERR_MESH, "tet index out of range");
  const int node_off[2] = {0, nn}, tet_off[2] = {0, nt};
  return create_any(1, node_off, tet_off, nodes, rest, mask, tets, &p->youngs, &p->nu, p, out);
}

static int create_batch_any(int ninst, const int *inst_nodes, const int *inst_tets, const float *nodes,
                            const float *rest, const int *mask, const int *tets, const float *youngs, const float *nu,
                            const int *inst_tiles, const pbs_params *p, pbs_solver **out) {
  if (!out) return set_error(PBS_ERR_ARG, "null out");
  *out = nullptr;
  if (ninst <= 0 || !inst_nodes || !inst_tets || !rest || !youngs || !nu)
    return set_error(PBS_ERR_ARG, "pbs_create_batch: bad arguments");
  int rc = validate_params(p);
  if (rc) return rc;
  std::vector<int> node_off(ninst + 1, 0), tet_off(ninst + 1, 0);
  for (int i = 0; i < ninst; ++i) {
    if (inst_nodes[i] <= 0 || inst_tets[i] < 0) return set_error(PBS_ERR_ARG, "instance sizes must be positive");
    if (!(youngs[i] > 0) || !(nu[i] > -1.0f && nu[i] < 0.5f)) return set_error(PBS_ERR_ARG, "bad instance material");
    if ((long)node_off[i] + inst_nodes[i] > 2147483647L || (long)tet_off[i] + inst_tets[i] > 2147483647L)
      return set_error(PBS_ERR_ARG, "batch too large");
    node_off[i + 1] = node_off[i] + inst_nodes[i];
    tet_off[i + 1] = tet_off[i] + inst_tets[i];
  }
  if (tet_off[ninst] > 0 && !tets) return set_error(PBS_ERR_ARG, "null tets");
  for (int i = 0; i < ninst; ++i)
    for (long k = 4L * tet_off[i]; k < 4L * tet_off[i + 1]; ++k)
      if (tets[k] < node_off[i] || tets[k] >= node_off[i + 1])
        return set_error(PBS_ERR_MESH, "tet index outside its instance's node range");
  if (inst_tiles) {
    for (int i = 0; i < ninst; ++i)
      if (inst_tiles[i] < 1 || inst_tiles[i] > 16383) return set_error(PBS_ERR_ARG, "instance tile size out of range");
    return create_impl(ninst, node_off.data(), tet_off.data(), nodes, rest, mask, tets, youngs, nu, p, 0, inst_tiles,
                       out);
  }
  return create_any(ninst, node_off.data(), tet_off.data(), nodes, rest, mask, tets, youngs, nu, p, out);
}

extern "C" int pbs_create_batch(int ninst, const int *inst_nodes, const int *inst_tets, const float *nodes,
                                const float *rest, const int *mask, const int *tets, const float *youngs,
                                const float *nu, const pbs_params *p, pbs_solver **out) {
  return create_batch_any(ninst, inst_nodes, inst_tets, nodes, rest, mask, tets, youngs, nu, nullptr, p, out);
}

extern "C" int pbs_create_batch_tiled(int ninst, const int *inst_nodes, const int *inst_tets, const float *nodes,
                                      const float *rest, const int *mask, const int *tets, const float *youngs,
                                      const float *nu, const int *inst_tile_tets, const pbs_params *p,
                                      pbs_solver **out) {
  if (!inst_tile_tets) return set_error(PBS_ERR_ARG, "null inst_tile_tets");
  return create_batch_any(ninst, inst_nodes, inst_tets, nodes, rest, mask, tets, youngs, nu, inst_tile_tets, p, out);
}

extern "C" int pbs_num_instances(const pbs_solver *s) {
  if (!s) return set_error(PBS_ERR_ARG, "null handle");
  return s->ninst;
}

extern "C" int pbs_instance_stats(const pbs_solver *s, int inst, pbs_step_stats *st) {
  if (!s || !st) return set_error(PBS_ERR_ARG, "null argument");
  if (inst < 0 || inst >= s->ninst) return set_error(PBS_ERR_ARG, "instance out of range");
  const Ctrl &c = s->ctrl_h[inst];
  *st = pbs_step_stats{};
  st->passes = c.passes;
  st->iterations = c.k;
  st->rollbacks = c.rollbacks;
  st->converged = c.converged;
  st->used_learned = c.used_learned;
  st->alpha = c.alpha;
  st->energy = c.E_old;
  st->g0_norm = c.g0;
  st->g_norm = c.g;
  return PBS_OK;
}

extern "C" int pbs_instance_tile_tets(const pbs_solver *s, int *tile_tets) {
  if (!s || !tile_tets) return set_error(PBS_ERR_ARG, "null argument");
  for (int i = 0; i < s->ninst; ++i) tile_tets[i] = s->plan.inst_tile_tets[i];
  return PBS_OK;
}

extern "C" int pbs_autotune_report(const pbs_solver *s, int max, int *tile_tets, double *us_per_pass, int *count) {
  if (!s || !count) return set_error(PBS_ERR_ARG, "null argument");
  const int n = (int)s->tune_tiles.size();
  for (int i = 0; i < std::min(n, max); ++i) {
    if (tile_tets) tile_tets[i] = s->tune_tiles[i];
    if (us_per_pass) us_per_pass[i] = s->tune_us[i];
  }
  *count = n;
  return PBS_OK;
}

extern "C" void pbs_destroy(pbs_solver *s) {
  if (!s) return;
  free_all(s);
  delete s;
}

extern "C" int pbs_sizes(const pbs_solver *s, int *nn, int *nt) {
  if (!s) return set_error(PBS_ERR_ARG, "null handle");
  if (nn) *nn = s->nn;
  if (nt) *nt = s->nt;
  return PBS_OK;
}

extern "C" int pbs_get_params(const pbs_solver *s, pbs_params *p) {
  if (!s || !p) return set_error(PBS_ERR_ARG, "null argument");
  *p = s->p;
  return PBS_OK;
}

extern "C" int pbs_set_params(pbs_solver *s, const pbs_params *p) {
  if (!s) return set_error(PBS_ERR_ARG, "null handle");
  int rc = validate_params(p);
  if (rc) return rc;
  if (p->precision != s->p.precision) return set_error(PBS_ERR_STATE, "precision is fixed at creation");
  const int tile = s->p.tile_tets;  // the tile plan is fixed at creation
  s->p = *p;
  s->p.tile_tets = tile;
  make_devparams(s);
  if (s->ninst == 1) {  // a batched solver keeps its per-instance materials
    s->inst_youngs[0] = p->youngs;
    s->inst_nu[0] = p->nu;
    rc = upload_lame(s);
    if (rc) return rc;
  }
  if (s->p.precision == PBS_FP64) {
    Launch<double>::setup_nodes(s->m64, s->dp, 256, s->st);
  } else {
    Launch<float>::setup_nodes(s->m32, s->dp, 256, s->st);
  }
  implicit_params_changed(s->implicit);  // the projective-dynamics matrix is rebuilt
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->st));
  return PBS_OK;
}

extern "C" int pbs_reset(pbs_solver *s) {
  if (!s) return set_error(PBS_ERR_ARG, "null handle");
  return s->p.precision == PBS_FP64 ? reset_impl<double>(s) : reset_impl<float>(s);
}

extern "C" int pbs_set_positions(pbs_solver *s, const float *xyz) {
  if (!s || !xyz) return set_error(PBS_ERR_ARG, "null argument");
  return s->p.precision == PBS_FP64 ? set_positions_impl<double>(s, xyz) : set_positions_impl<float>(s, xyz);
}

extern "C" int pbs_get_positions(pbs_solver *s, float *xyz) {
  if (!s || !xyz) return set_error(PBS_ERR_ARG, "null argument");
  return s->p.precision == PBS_FP64 ? download_planes<double>(s, s->m64.xpos, xyz) : download_planes<float>(s, s->m32.xpos, xyz);
}

template <typename R> static int get_pos64_impl(pbs_solver *s, double *xyz) {
  Launch<R>::planes_to_aos64(mesh<R>(s).xpos, s->stage64, mesh<R>(s).perm, s->nn, s->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(xyz, s->stage64, sizeof(double) * 3 * (size_t)s->nn, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return PBS_OK;
}

extern "C" int pbs_get_positions_f64(pbs_solver *s, double *xyz) {
  if (!s || !xyz) return set_error(PBS_ERR_ARG, "null argument");
  return s->p.precision == PBS_FP64 ? get_pos64_impl<double>(s, xyz) : get_pos64_impl<float>(s, xyz);
}

template <typename R> static int set_vel_impl(pbs_solver *s, const float *v, const float *vl) {
  int rc = 0;
  if (v) rc = upload_planes<R>(s, v, mesh<R>(s).v);
  if (rc) return rc;
  if (vl) rc = upload_planes<R>(s, vl, mesh<R>(s).vl);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s->st));
  return PBS_OK;
}
extern "C" int pbs_set_velocities(pbs_solver *s, const float *v, const float *vl) {
  if (!s) return set_error(PBS_ERR_ARG, "null handle");
  return s->p.precision == PBS_FP64 ? set_vel_impl<double>(s, v, vl) : set_vel_impl<float>(s, v, vl);
}
template <typename R> static int get_vel_impl(pbs_solver *s, float *v, float *vl) {
  int rc = 0;
  if (v) rc = download_planes<R>(s, mesh<R>(s).v, v);
  if (rc) return rc;
  if (vl) rc = download_planes<R>(s, mesh<R>(s).vl, vl);
  return rc;
}
extern "C" int pbs_get_velocities(pbs_solver *s, float *v, float *vl) {
  if (!s) return set_error(PBS_ERR_ARG, "null handle");
  return s->p.precision == PBS_FP64 ? get_vel_impl<double>(s, v, vl) : get_vel_impl<float>(s, v, vl);
}

// ---- state snapshot ----
// File: "PBSSTAT1", int32 precision (4 or 8 bytes per value), int32 node count, int32 instance
// count, int32 0, int32 node count of each instance, then x, v and v_last, each 3N values AoS in
// the caller's node order at the solver's precision (so an FP64 run resumes bitwise).
static const char kStateMagic[8] = {'P', 'B', 'S', 'S', 'T', 'A', 'T', '1'};

template <typename R> static int state_save_impl(pbs_solver *s, FILE *f) {
  DevMesh<R> &m = mesh<R>(s);
  const size_t n3 = 3 * (size_t)s->nn;
  std::vector<R> h(n3);
  const R *src[3] = {m.xpos, m.v, m.vl};
  for (const R *planes : src) {
    if (sizeof(R) == 8) {
      Launch<R>::planes_to_aos64(planes, s->stage64, m.perm, s->nn, s->st);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(h.data(), s->stage64, sizeof(double) * n3, hipMemcpyDeviceToHost, s->st));
    } else {
      Launch<R>::planes_to_aos(planes, s->stage, m.perm, s->nn, s->st);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(h.data(), s->stage, sizeof(float) * n3, hipMemcpyDeviceToHost, s->st));
    }
    HIPCHK(hipStreamSynchronize(s->st));
    if (std::fwrite(h.data(), sizeof(R), n3, f) != n3) return set_error(PBS_ERR_ARG, "state snapshot: write failed");
  }
  return PBS_OK;
}

template <typename R> static int state_load_impl(pbs_solver *s, FILE *f) {
  DevMesh<R> &m = mesh<R>(s);
  const size_t n3 = 3 * (size_t)s->nn;
  std::vector<R> h[3];
  for (auto &a : h) {  // read everything before touching the device state
    a.resize(n3);
    if (std::fread(a.data(), sizeof(R), n3, f) != n3) return set_error(PBS_ERR_ARG, "state snapshot: file too short");
  }
  R *dst[3] = {m.xpos, m.v, m.vl};
  for (int k = 0; k < 3; ++k) {
    if (sizeof(R) == 8) {
      HIPCHK(hipMemcpyAsync(s->stage64, h[k].data(), sizeof(double) * n3, hipMemcpyHostToDevice, s->st));
      Launch<R>::aos64_to_planes(s->stage64, dst[k], m.perm, s->nn, s->st);
    } else {
      HIPCHK(hipMemcpyAsync(s->stage, h[k].data(), sizeof(float) * n3, hipMemcpyHostToDevice, s->st));
      Launch<R>::aos_to_planes(s->stage, dst[k], m.perm, s->nn, s->st);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->st));
  }
  return PBS_OK;
}

extern "C" int pbs_state_save(pbs_solver *s, const char *path) {
  if (!s || !path) return set_error(PBS_ERR_ARG, "null argument");
  FILE *f = std::fopen(path, "wb");
  if (!f) return set_error(PBS_ERR_ARG, std::string("state snapshot: cannot open ") + path);
  const bool f64 = s->p.precision == PBS_FP64;
  const int32_t head[4] = {f64 ? 8 : 4, s->nn, s->ninst, 0};
  std::vector<int32_t> inst(s->ninst);
  for (int i = 0; i < s->ninst; ++i) inst[i] = s->plan.inst_node_off[i + 1] - s->plan.inst_node_off[i];
  int rc = PBS_OK;
  if (std::fwrite(kStateMagic, 1, 8, f) != 8 || std::fwrite(head, sizeof(int32_t), 4, f) != 4 ||
      std::fwrite(inst.data(), sizeof(int32_t), inst.size(), f) != inst.size())
    rc = set_error(PBS_ERR_ARG, "state snapshot: write failed");
  if (!rc) rc = f64 ? state_save_impl<double>(s, f) : state_save_impl<float>(s, f);
  if (std::fclose(f) != 0 && !rc) rc = set_error(PBS_ERR_ARG, "state snapshot: write failed");
  return rc;
}

extern "C" int pbs_state_load(pbs_solver *s, const char *path) {
  if (!s || !path) return set_error(PBS_ERR_ARG, "null argument");
  FILE *f = std::fopen(path, "rb");
  if (!f) return set_error(PBS_ERR_ARG, std::string("state snapshot: cannot open ") + path);
  const bool f64 = s->p.precision == PBS_FP64;
  char magic[8];
  int32_t head[4];
  int rc = PBS_OK;
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kStateMagic, 8) != 0 ||
      std::fread(head, sizeof(int32_t), 4, f) != 4)
    rc = set_error(PBS_ERR_ARG, "state snapshot: not a snapshot file");
  else if (head[0] != (f64 ? 8 : 4) || head[1] != s->nn || head[2] != s->ninst)
    rc = set_error(PBS_ERR_ARG, "state snapshot: precision, node count or instance count differs from this solver's");
  if (!rc) {
    std::vector<int32_t> inst(s->ninst);
    if (std::fread(inst.data(), sizeof(int32_t), inst.size(), f) != inst.size())
      rc = set_error(PBS_ERR_ARG, "state snapshot: file too short");
    for (int i = 0; i < s->ninst && !rc; ++i)
      if (inst[i] != s->plan.inst_node_off[i + 1] - s->plan.inst_node_off[i])
        rc = set_error(PBS_ERR_ARG, "state snapshot: instance node counts differ from this solver's");
  }
  if (!rc) rc = f64 ? state_load_impl<double>(s, f) : state_load_impl<float>(s, f);
  std::fclose(f);
  return rc;
}

// ---- the step ----
// One descent pass; timed: each kernel's launch records its own start and end (the dispatch's
// timestamps, as a kernel trace reports them).
template <typename R> static void launch_pass(pbs_solver *s, bool timed) {
  auto ev = [&](int i, int k) -> hipEvent_t { return i >= 0 ? s->pool[i + k] : nullptr; };
  int i = timed ? timed_slot(s, s->dp.f32math ? K_TILE_MIXED : K_TILE) : -1;
  Launch<R>::tile_pass(mesh<R>(s), s->td, s->dp, s->st, ev(i, 0), ev(i, 1));
  i = timed ? timed_slot(s, K_NODE) : -1;
  Launch<R>::node_pass(mesh<R>(s), s->td, s->dp, s->st, ev(i, 0), ev(i, 1));
  i = timed ? timed_slot(s, K_CONTROL) : -1;
  Launch<R>::control(mesh<R>(s), s->td, s->dp, s->st, ev(i, 0), ev(i, 1));
}

// Captures G descent passes into a HIP graph, cached per arithmetic (FP64 or mixed-mode FP32).
// Timed replays are launched eagerly (launch_pass with timing), not from a graph.
template <typename R> static int graph_build(pbs_solver *s, int G) {
  auto &gs = s->graph[s->dp.f32math ? 1 : 0];
  if (gs.exec) (void)hipGraphExecDestroy(gs.exec);
  gs.exec = nullptr;
  gs.passes = 0;
  HIPCHK(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
  hipError_t err = hipSuccess;
  for (int j = 0; j < G && err == hipSuccess; ++j) {
    Launch<R>::tile_pass(mesh<R>(s), s->td, s->dp, s->st);
    Launch<R>::node_pass(mesh<R>(s), s->td, s->dp, s->st);
    Launch<R>::control(mesh<R>(s), s->td, s->dp, s->st);
    err = hipGetLastError();
  }
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(s->st, &g);
  if (err == hipSuccess) err = e2;
  if (err == hipSuccess) err = hipGraphInstantiate(&gs.exec, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  if (err != hipSuccess) {
    gs.exec = nullptr;
    return set_error(PBS_ERR_HIP, std::string("descent graph capture: ") + hipGetErrorString(err));
  }
  gs.dp = s->dp;
  gs.passes = G;
  return PBS_OK;
}

template <typename R> static int graph_ready(pbs_solver *s, int G) {
  const auto &gs = s->graph[s->dp.f32math ? 1 : 0];
  if (gs.exec && gs.passes == G && std::memcmp(&gs.dp, &s->dp, sizeof(DevParams)) == 0) return PBS_OK;
  return graph_build<R>(s, G);
}

template <typename R> static int step_impl(pbs_solver *s, const float *ext, float *out, pbs_step_stats *stats) {
  DevMesh<R> &m = mesh<R>(s);
  if (ext) {
    HIPCHK(hipMemcpyAsync(s->stage, ext, sizeof(float) * 3 * (size_t)s->nn, hipMemcpyHostToDevice, s->st));
    aos_to_planes_f(s->stage, m.ext, m.perm, s->nn, s->st);
    s->dp.has_ext = 1;
  } else {
    s->dp.has_ext = 0;
  }
  // descent passes go through captured graphs unless graph_passes < 0
  const int G = s->p.graph_passes == 0 ? 32 : s->p.graph_passes;
  const bool use_graph = G > 0;
  const int period = s->prof_period;
  s->dp.f32math = 0;  // step start, warm start and eager passes: FP64 arithmetic
  HIPCHK(hipEventRecord(s->ev0, s->st));
  s->pass_ordinal = -1;
  int i = timed_begin(s, K_INIT);
  Launch<R>::step_init(m, s->td, s->dp, s->bs, s->st);
  timed_end(s, i);
  HIPCHK(hipGetLastError());
  if (s->learn) {  // features / learned guess / selection by gradient norm
    i = timed_begin(s, K_LEARN);
    int rc = learn_before_descent<R>(s);
    timed_end(s, i);
    if (rc) return rc;
  }
  const bool relg = s->dp.relgrad_tol > 0;
  const long cap = relg ? (long)s->p.max_iterations + 256 : (long)s->p.iterations + 256;
  long launched = 0;
  long n = relg ? 32 : std::max(s->p.iterations, 1);
  // control blocks of all instances: the loop runs while any instance is active; a pass counts
  // as executed when some instance ran it
  std::vector<Ctrl> &ch = s->ctrl_h;
  auto read_ctrl = [&]() -> int {
    HIPCHK(hipMemcpyAsync(ch.data(), m.ctrl, sizeof(Ctrl) * ch.size(), hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return PBS_OK;
  };
  auto max_passes = [&]() {
    int mp = 0;
    for (const Ctrl &q : ch) mp = std::max(mp, q.passes);
    return mp;
  };
  auto all_done = [&]() {
    for (const Ctrl &q : ch)
      if (!q.done) return false;
    return true;
  };
  // mixed mode: FP32 per-tet arithmetic until the last replay's |g|/|g0| (largest over the active
  // instances) is at or below mixed_switch, or some instance's |g| is within kMixedNoise of its
  // internal-force norm |f| (the FP32 forces' rounding, ~5e-5 |f| measured, would then dominate
  // g), FP64 from then on.  A rollback during an FP32 replay means the FP32 forces no longer
  // resolve the remaining gradient: the step goes on in FP64 from the restored iterate, with the
  // step length the step started with.
  constexpr double kMixedNoise = 2e-4;
  const bool mixed = use_graph && relg && s->p.mixed_switch > 0 && sizeof(R) == 8;
  bool fp64_phase = !mixed;
  std::vector<float> alpha0;
  std::vector<int> rb0;
  if (mixed) {
    int rc = read_ctrl();
    if (rc) return rc;
    for (const Ctrl &q : ch) {
      alpha0.push_back(q.alpha);
      rb0.push_back(q.rollbacks);
    }
  }
  while (true) {
    if (use_graph) {
      if (!fp64_phase) {
        double worst = -1.0;
        bool known = true;
        bool noisy = false;
        for (const Ctrl &q : ch)
          if (!q.done) {
            if (!(q.g0 > 0.0)) known = false;
            else worst = std::max(worst, q.g / q.g0);
            if (q.g <= kMixedNoise * q.fnorm) noisy = true;
          }
        if (launched > 0 && ((known && worst <= s->p.mixed_switch) || noisy)) fp64_phase = true;
      }
      s->dp.f32math = fp64_phase ? 0 : 1;
      int rc = graph_ready<R>(s, G);
      if (rc) return rc;
      const int gi = s->dp.f32math ? 1 : 0;
      const long reps = (n + G - 1) / G;
      for (long r = 0; r < reps; ++r) {
        const bool timed = period && (s->prof_tick++ % period) == 0;
        if (timed) {  // the same G passes launched eagerly, each kernel timing itself
          for (long j = 0; j < G; ++j) {
            s->pass_ordinal = launched + j;
            launch_pass<R>(s, true);
          }
          HIPCHK(hipGetLastError());
          int rc = read_ctrl();  // synchronises: the events are complete and can be reused
          if (rc) return rc;
          timed_collect(s, max_passes());
        } else {
          HIPCHK(hipGraphLaunch(s->graph[gi].exec, s->st));
        }
        launched += G;
      }
    } else {
      for (long k = 0; k < n; ++k) {
        s->pass_ordinal = launched + k;
        launch_pass<R>(s, period && (s->prof_tick++ % period) == 0);
      }
      HIPCHK(hipGetLastError());
      launched += n;
    }
    int rc = read_ctrl();
    if (rc) return rc;
    if (!fp64_phase) {
      bool rolled = false;
      for (size_t q = 0; q < ch.size(); ++q)
        if (ch[q].rollbacks > rb0[q]) {
          ch[q].alpha = alpha0[q];
          rolled = true;
        }
      if (rolled) {
        fp64_phase = true;
        HIPCHK(hipMemcpyAsync(m.ctrl, ch.data(), sizeof(Ctrl) * ch.size(), hipMemcpyHostToDevice, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
      }
    }
    if (all_done()) break;
    if (launched >= cap) return set_error(PBS_ERR_STATE, "descent loop did not terminate within the pass cap");
    n = 1;
    for (const Ctrl &q : ch)
      if (!q.done) n = std::max<long>(n, relg ? 32 : std::max(1, q.budget - q.k) + 1);
  }
  s->dp.f32math = 0;
  s->pass_ordinal = -1;
  i = timed_begin(s, K_FINAL);
  Launch<R>::step_final(m, s->td, s->dp, s->st);
  timed_end(s, i);
  HIPCHK(hipGetLastError());
  if (s->learn) {
    int rc = learn_after_step<R>(s);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(s->ev1, s->st));
  if (out) {
    Launch<R>::planes_to_aos(m.xpos, s->stage, m.perm, s->nn, s->st);
    HIPCHK(hipMemcpyAsync(out, s->stage, sizeof(float) * 3 * (size_t)s->nn, hipMemcpyDeviceToHost, s->st));
  }
  HIPCHK(hipStreamSynchronize(s->st));
  if (period) timed_collect(s, max_passes());
  if (stats) {
    // one instance: its loop statistics; several: passes = the most any instance ran, the counts
    // summed, alpha the smallest, energy summed, gradient norms combined as sqrt(sum of squares)
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, s->ev0, s->ev1);
    const Ctrl &c = ch[0];
    *stats = pbs_step_stats{};
    stats->passes = max_passes();
    stats->alpha = c.alpha;
    double g0 = 0.0, g = 0.0;
    for (const Ctrl &q : ch) {
      stats->iterations += q.k;
      stats->rollbacks += q.rollbacks;
      stats->converged += q.converged;
      stats->used_learned += q.used_learned;
      stats->alpha = std::min(stats->alpha, q.alpha);
      stats->energy += q.E_old;
      g0 += q.g0 * q.g0;
      g += q.g * q.g;
    }
    stats->g0_norm = s->ninst == 1 ? c.g0 : std::sqrt(g0);
    stats->g_norm = s->ninst == 1 ? c.g : std::sqrt(g);
    stats->gpu_ms = ms;
    stats->g_learned = c.used_learned || (s->learn && s->learn->enabled) ? std::sqrt(c.g_learned) : 0.0;
  }
  return PBS_OK;
}

extern "C" int pbs_step(pbs_solver *s, const float *ext, float *out, pbs_step_stats *st) {
  if (!s) return set_error(PBS_ERR_ARG, "null handle");
  return s->p.precision == PBS_FP64 ? step_impl<double>(s, ext, out, st) : step_impl<float>(s, ext, out, st);
}

extern "C" int pbs_do_descent_opt(pbs_solver *s, float h, int iterations, const float *ext, float *out) {
  (void)h;  // ignored as in the reference (CUDAInterface.cu:911 uses m_timestep)
  if (!s || !ext || !out || iterations < 0) return set_error(PBS_ERR_ARG, "pbs_do_descent_opt: bad arguments");
  int saved = s->p.iterations;
  double saved_tol = s->p.relgrad_tol;
  s->p.iterations = iterations;
  s->p.relgrad_tol = 0.0;
  make_devparams(s);
  int rc = pbs_step(s, ext, out, nullptr);
  s->p.iterations = saved;
  s->p.relgrad_tol = saved_tol;
  make_devparams(s);
  return rc;
}

// ---- material calls ----
// Per-tet records (width values per tet, solver tet order on the device) -> caller tet order.
template <typename T> static int download_records(pbs_solver *s, const T *dev, float *host, int width) {
  const size_t n = (size_t)width * s->nt;
  std::vector<T> tmp(n);
  HIPCHK(hipMemcpyAsync(tmp.data(), dev, sizeof(T) * n, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  for (int t = 0; t < s->nt; ++t) {
    const size_t o = (size_t)width * s->plan.tperm[t];
    for (int k = 0; k < width; ++k) host[o + k] = (float)tmp[(size_t)width * t + k];
  }
  return PBS_OK;
}

template <typename R> static int forces_impl(pbs_solver *s, const float *pos, float *tf, float *nf, float *te) {
  DevMesh<R> &m = mesh<R>(s);
  int rc = upload_planes<R>(s, pos, m.ring.b[0]);
  if (rc) return rc;
  Launch<R>::standalone_ctrl(m, 0, s->st);
  Launch<R>::tet(m, s->td, s->dp, s->bs, 0, 1, s->st);
  HIPCHK(hipGetLastError());
  if (nf) {
    Launch<R>::gather(m, s->dp, s->bs, 0, m.ring.b[1], s->st);
    rc = download_planes<R>(s, m.ring.b[1], nf);
    if (rc) return rc;
  }
  if (tf && s->nt > 0) {
    rc = download_records<R>(s, m.frec, tf, 12);
    if (rc) return rc;
  }
  if (te && s->nt > 0) {
    rc = download_records<R>(s, m.tete, te, 1);
    if (rc) return rc;
  }
  HIPCHK(hipStreamSynchronize(s->st));
  return PBS_OK;
}

extern "C" int pbs_compute_forces(pbs_solver *s, const float *pos, float *tf, float *nf, float *te) {
  if (!s || !pos) return set_error(PBS_ERR_ARG, "null argument");
  return s->p.precision == PBS_FP64 ? forces_impl<double>(s, pos, tf, nf, te) : forces_impl<float>(s, pos, tf, nf, te);
}

template <typename R> static int hd_impl(pbs_solver *s, const float *pos, float *nh, float *th) {
  DevMesh<R> &m = mesh<R>(s);
  int rc = upload_planes<R>(s, pos, m.ring.b[0]);
  if (rc) return rc;
  Launch<R>::standalone_ctrl(m, 0, s->st);
  Launch<R>::tet(m, s->td, s->dp, s->bs, 1, 0, s->st);
  HIPCHK(hipGetLastError());
  if (nh) {
    Launch<R>::gather(m, s->dp, s->bs, 1, m.hd, s->st);
    planes_to_aos_f(m.hd, s->stage, m.perm, s->nn, s->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(nh, s->stage, sizeof(float) * 3 * (size_t)s->nn, hipMemcpyDeviceToHost, s->st));
  }
  HIPCHK(hipStreamSynchronize(s->st));
  if (th && s->nt > 0) {
    rc = download_records<float>(s, m.hrec, th, 12);
    if (rc) return rc;
  }
  return PBS_OK;
}

extern "C" int pbs_compute_hessian_diag(pbs_solver *s, const float *pos, float *nh, float *th) {
  if (!s || !pos) return set_error(PBS_ERR_ARG, "null argument");
  return s->p.precision == PBS_FP64 ? hd_impl<double>(s, pos, nh, th) : hd_impl<float>(s, pos, nh, th);
}

template <typename R> static int tile_forces_impl(pbs_solver *s, const float *pos, double *nf, double *nh, double *en) {
  DevMesh<R> &m = mesh<R>(s);
  int rc = upload_planes<R>(s, pos, m.ring.b[0]);
  if (rc) return rc;
  DevParams dp = s->dp;
  dp.f32math = 0;
  Launch<R>::standalone_ctrl(m, 0, s->st, nh ? 1 : 0, 1);
  Launch<R>::tile_pass(m, s->td, dp, s->st);
  HIPCHK(hipGetLastError());
  const size_t n3 = 3 * (size_t)s->nn;
  for (int which = 0; which < 2; ++which) {
    double *out = which == 0 ? nf : nh;
    if (!out) continue;
    Launch<R>::sum_partials(m, s->td, dp, which, s->stage64, s->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, s->stage64, sizeof(double) * n3, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
  }
  if (en) {
    std::vector<double> tp(std::max(s->td.ntiles, 1));
    HIPCHK(hipMemcpyAsync(tp.data(), s->td.tpart, sizeof(double) * s->td.ntiles, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    double e = 0.0;
    for (int b = 0; b < s->td.ntiles; ++b) e += tp[b];
    *en = e;
  }
  HIPCHK(hipStreamSynchronize(s->st));
  return PBS_OK;
}

extern "C" int pbs_tile_forces(pbs_solver *s, const float *pos, double *node_forces, double *node_hd, double *energy) {
  if (!s || !pos) return set_error(PBS_ERR_ARG, "null argument");
  if (s->ninst != 1) return set_error(PBS_ERR_STATE, "pbs_tile_forces: single-mesh handles only");
  return s->p.precision == PBS_FP64 ? tile_forces_impl<double>(s, pos, node_forces, node_hd, energy)
                                    : tile_forces_impl<float>(s, pos, node_forces, node_hd, energy);
}

template <typename R> static int rest_data_impl(pbs_solver *s, float *dminv, float *vol, float *mass, float *gforce) {
  DevMesh<R> &m = mesh<R>(s);
  const TilePlan &P = s->plan;
  const size_t N = s->nn, T = s->nt;
  if (dminv && T > 0) {
    std::vector<float> planes(9 * T);
    HIPCHK(hipMemcpyAsync(planes.data(), m.dmi, sizeof(float) * 9 * T, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    for (size_t t = 0; t < T; ++t)
      for (int e = 0; e < 9; ++e) dminv[9 * (size_t)P.tperm[t] + e] = planes[e * T + t];
  }
  if (vol && T > 0) {
    std::vector<float> v(T);
    HIPCHK(hipMemcpyAsync(v.data(), m.vol, sizeof(float) * T, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    for (size_t t = 0; t < T; ++t) vol[P.tperm[t]] = v[t];
  }
  if (mass) {
    std::vector<float> v(N);
    HIPCHK(hipMemcpyAsync(v.data(), m.mass, sizeof(float) * N, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    for (size_t i = 0; i < N; ++i) mass[P.perm[i]] = v[i];
  }
  if (gforce) {
    planes_to_aos_f(m.gforce, s->stage, m.perm, s->nn, s->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(gforce, s->stage, sizeof(float) * 3 * N, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
  }
  return PBS_OK;
}

extern "C" int pbs_get_rest_data(pbs_solver *s, float *dminv, float *vol, float *mass, float *gforce) {
  if (!s) return set_error(PBS_ERR_ARG, "null handle");
  return s->p.precision == PBS_FP64 ? rest_data_impl<double>(s, dminv, vol, mass, gforce)
                                    : rest_data_impl<float>(s, dminv, vol, mass, gforce);
}

// ---- measurement ----
extern "C" int pbs_set_profiling(pbs_solver *s, int period) {
  if (!s) return set_error(PBS_ERR_ARG, "null handle");
  if (period < 0) return set_error(PBS_ERR_ARG, "profiling period must be >= 0");
  s->prof_period = period;
  s->prof_tick = 0;
  return PBS_OK;
}
extern "C" int pbs_reset_kernel_times(pbs_solver *s) {
  if (!s) return set_error(PBS_ERR_ARG, "null handle");
  for (int i = 0; i < K_COUNT; ++i) {
    s->kt_ms[i] = 0;
    s->kt_n[i] = 0;
  }
  return PBS_OK;
}
extern "C" int pbs_kernel_times(pbs_solver *s, int max, const char **names, double *avg_ms, long *launches, int *count) {
  if (!s || !count) return set_error(PBS_ERR_ARG, "null argument");
  int c = 0;
  for (int i = 0; i < K_COUNT && c < max; ++i, ++c) {
    if (names) names[c] = kKernelNames[i];
    if (avg_ms) avg_ms[c] = s->kt_n[i] ? s->kt_ms[i] / (double)s->kt_n[i] : 0.0;
    if (launches) launches[c] = s->kt_n[i];
  }
  *count = c;
  return PBS_OK;
}
