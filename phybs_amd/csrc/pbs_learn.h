// pbs_learn.h -- host-side state of the learned warm start (kernels in pbs_learn.hip).
#pragma once
#include <vector>

#include "pbs_internal.h"

struct pbs_solver;

namespace pbs {

struct LearnState {
  // network: FP32 master parameters, Adam moments, last gradient, BF16 MFMA arrangement
  float *params = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr;
  unsigned short *wA = nullptr;
  float scales[MLP_NSCALE] = {1, 1, 1, 1, 1, 1, 1};
  bool have_model = false;
  long adam_t = 0;
  bool enabled = false;  // pbs_step starts from the learned guess when it is not worse
  int smooth_iters = 0;  // averaging sweeps applied to the learned correction
  std::vector<float> betas = {1.0f};  // steps along the correction tried by the gradient test
  // node adjacency (solver order) and the rank of each free node among the free nodes
  int *adj_off = nullptr, *adj = nullptr, *free_rank = nullptr;
  long nfree = 0;
  // per-step buffers: features (N x 18) and the inertial target y (3 planes of R)
  float *feat = nullptr;
  float *smooth = nullptr;  // 2 x N float4: the correction and its next sweep
  int *sell_off = nullptr, *sell_col = nullptr;  // adjacency as SELL-64 slices (smoothing sweeps)
  void *y = nullptr;
  // collected samples
  bool collecting = false;
  long cap = 0, count = 0;
  int stride = 1;       // each step keeps the free nodes of rank r with r % stride == phase
  long steps_seen = 0;  // collected steps (phase = steps_seen % stride)
  float *ds_x = nullptr, *ds_y = nullptr;
  // training scratch
  float *grad_part = nullptr, *loss_part = nullptr;
  int part_cap = 0;
  double *loss_acc = nullptr;
  int *idx = nullptr;
  long idx_cap = 0;
  float *io = nullptr;  // staging for the standalone calls
  long io_cap = 0;
  std::vector<void *> allocs;
};

MlpDev mlp_dev(const LearnState &L);
// Called by pbs_step: before the descent (after the step-start kernel) and after it.
template <typename R> int learn_before_descent(pbs_solver *s);
template <typename R> int learn_after_step(pbs_solver *s);

}  // namespace pbs
