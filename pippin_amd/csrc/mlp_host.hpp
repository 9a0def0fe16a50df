// Host-side launchers of the learned warm start (mlp.hip), used by the C ABI
// in runtime.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "dmodel.hpp"

namespace phm {

// The step kernel with the learned warm start for K envs per wavefront (spec:
// the model-specialised instance for SpecAnt).
const void* step_kernel(int K, int spec);
void launch_step(int K, int spec, unsigned grid, size_t lds, hipStream_t st, const phk::DModel* M,
                 const phk::DState* S, int64_t begin, int64_t end, int64_t nstep, const double* ctrl,
                 double* qpos_out);

// One network and its optimiser state, in device memory.
struct Net {
  int h = 0;
  float* P = nullptr;      // FP32 parameters W1 b1 W2 b2 W3 b3 W4 b4 (mlp_woff / mlp_boff)
  float* m = nullptr;      // Adam first moments
  float* v = nullptr;      // Adam second moments
  float* G = nullptr;      // gradient of the last batch
  uint16_t* wbf = nullptr;   // BF16 W_l, [IN_l][OUT_l], packed by mlp_wbf_off
  uint16_t* wtbf = nullptr;  // BF16 W_l^T, [OUT_l][IN_l]
};

// Scratch for one batch of up to bmax samples (a multiple of 16).
struct Batch {
  int64_t bmax = 0;
  int splits = 0;           // split-K partitions of the weight gradient
  uint16_t* x0 = nullptr;   // bmax x 32 inputs (BF16)
  float* z = nullptr;       // 3 x bmax x h pre-activations
  uint16_t* a = nullptr;    // 3 x bmax x h activations (BF16)
  uint16_t* dz = nullptr;   // 3 x bmax x h, then bmax x 16: loss gradients w.r.t. pre-activations (BF16)
  float* y = nullptr;       // bmax x 16 outputs
  float* loss = nullptr;    // bmax / 16 per-tile sums of squared errors
  float* gp = nullptr;      // splits x nparam partial gradients
};

phk::MlpW view(const Net& n, int mode);
// BF16 copies of the weights from the FP32 parameters.
void to_bf16(const Net& n, hipStream_t st);
// Forward (outputs in b.y, per-tile squared errors in b.loss when tgt is
// given) and, with backward, the mean-squared-error gradient into n.G. The
// batch is rows [0, B) of feat / tgt (identity) or B rows drawn from [0, nsrc)
// by a hash of (seed, iter, sample).
void forward_backward(const Net& n, const Batch& b, const float* feat, const float* tgt, int64_t nsrc, int64_t B,
                      bool identity, uint64_t seed, int64_t iter, bool backward, hipStream_t st);
// One Adam step (t counts from 1), then the BF16 copies.
void adam(const Net& n, int64_t t, float lr, hipStream_t st);
// Append the recorded contacts of envs [begin, end) to the ring of capacity
// cap at position head (env order, then contact order). d_scratch holds
// 1 + (end - begin) int64: d_scratch[0] receives the number appended.
void collect(const int* cnt, const float* feat, const float* tgt, int64_t begin, int64_t end, float* rfeat,
             float* rtgt, int64_t cap, int64_t head, int64_t* d_scratch, hipStream_t st);

}  // namespace phm
