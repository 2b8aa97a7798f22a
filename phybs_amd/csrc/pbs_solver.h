// pbs_solver.h -- the solver handle behind the C ABI, shared by the host translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "pbs_host.h"
#include "pbs_internal.h"
#include "pbs_tiles.h"
#include "phybs.h"

namespace pbs {
struct LearnState;  // pbs_learn.h
void learn_release(LearnState *L);
struct ImplicitState;  // pbs_implicit.h
void implicit_release(ImplicitState *S);
void implicit_params_changed(ImplicitState *S);  // masses or time step may have changed
}  // namespace pbs

using namespace pbs;

#define HIPCHK(expr)                                                                                  \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess) return set_error(PBS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

enum KernelId {
  K_TILE = 0, K_NODE, K_CONTROL, K_INIT, K_FINAL, K_LEARN, K_IE_ASSEMBLE, K_IE_SPMV, K_IE_UPDATE, K_PD_LOCAL, K_TILE_MIXED, K_COUNT
};
static const char *const kKernelNames[K_COUNT] = {"k_tile",        "k_node",    "k_control",   "k_step_init",
                                                  "k_step_final",  "k_warm_start", "k_ie_assemble", "k_ie_spmv",
                                                  "k_ie_update",   "k_pd_local",   "k_tile_mixed"};

struct pbs_solver {
  int nn = 0, nt = 0;
  int ninst = 1;                          // independent meshes (batched solver)
  std::vector<float> inst_youngs, inst_nu;
  std::vector<Ctrl> ctrl_h;               // last copy of the control blocks
  std::vector<int> tune_tiles;            // autotuner: candidate tile sizes ...
  std::vector<double> tune_us;            // ... and their time per descent pass (tile -2: per-instance sizes)
  std::vector<int> tune_class_tiles;      // autotuner, batches: fastest size of each instance class
  pbs_params p{};
  DevParams dp{};
  int bs = 256, ntb = 0, nnb = 0;
  hipStream_t st = nullptr;
  int device = 0;
  DevMesh<float> m32{};
  DevMesh<double> m64{};
  TileDev td{};
  TilePlan plan;
  size_t lds_budget = 65536;
  std::vector<float> rest_h;
  float *stage = nullptr;     // 3N float AoS staging
  double *stage64 = nullptr;  // 3N double AoS staging
  std::vector<void *> allocs;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<hipEvent_t> pool;
  struct Pending {
    int kid, ev;  // kernel id, first of its two events
    long pass;    // pass ordinal within the step (-1: not a descent pass)
  };
  std::vector<Pending> pending;
  int pool_used = 0;
  long pass_ordinal = -1;  // ordinal given to the descent-pass launches being recorded
  // HIP graphs of graph_passes descent passes, replayed by pbs_step: [0] FP64 arithmetic,
  // [1] mixed-mode FP32 arithmetic
  struct PassGraph {
    hipGraphExec_t exec = nullptr;
    DevParams dp{};  // parameters baked into the graph
    int passes = 0;
  } graph[2];
  int prof_period = 0;  // 0: no timing; n: time every n-th graph replay (or pass without graphs)
  long prof_tick = 0;
  double kt_ms[K_COUNT] = {0};
  long kt_n[K_COUNT] = {0};
  pbs::LearnState *learn = nullptr;  // learned warm start (created on first use)
  pbs::ImplicitState *implicit = nullptr;  // backward-Euler matrix and CG buffers (first use)
};

template <typename R> inline DevMesh<R> &mesh(pbs_solver *s);
template <> inline DevMesh<float> &mesh<float>(pbs_solver *s) { return s->m32; }
template <> inline DevMesh<double> &mesh<double>(pbs_solver *s) { return s->m64; }

template <typename T> inline int dalloc(pbs_solver *s, T **p, size_t n) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(PBS_ERR_HIP, std::string("hipMalloc(") + std::to_string(n * sizeof(T)) + " B): " + hipGetErrorString(e));
  }
  s->allocs.push_back(q);
  *p = (T *)q;
  return PBS_OK;
}

