// Matrix-core pieces of the learned warm start (device code).
//
// The network runs in transposed form, Z^T = W^T X^T, so that one MFMA's
// result layout is the next MFMA's B-operand layout: for
// v_mfma_f32_16x16x16_bf16, lane l = 16 q + j holds
//   A: row j, k = 4q .. 4q+3;  B: column j, k = 4q .. 4q+3;
//   D: column j, rows 4q .. 4q+3.
// With samples as columns, lane (j, q) feeds inputs 4q..4q+3 of sample j and
// receives outputs 4q..4q+3 of sample j, which after the activation are the
// B operand of the next layer's k-block. No LDS traffic between layers.
#pragma once
#include <hip/hip_runtime.h>

#include "dmodel.hpp"

namespace phk {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma_bf16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x4 ld_bf4(const uint16_t* p) { return *(const PH_AS_GLOBAL bf16x4*)p; }
__device__ __forceinline__ void st_bf4(uint16_t* p, bf16x4 v) { *(PH_AS_GLOBAL bf16x4*)p = v; }
__device__ __forceinline__ f32x4 ld_f4(const float* p) { return *(const PH_AS_GLOBAL f32x4*)p; }
__device__ __forceinline__ void st_f4(float* p, f32x4 v) { *(PH_AS_GLOBAL f32x4*)p = v; }
__device__ __forceinline__ bf16x4 pack_bf4(f32x4 v) {
  bf16x4 r;
  r[0] = (short)f2bf(v[0]);
  r[1] = (short)f2bf(v[1]);
  r[2] = (short)f2bf(v[2]);
  r[3] = (short)f2bf(v[3]);
  return r;
}
__device__ __forceinline__ f32x4 unpack_bf4(bf16x4 v) {
  f32x4 r;
  r[0] = bf2f((uint16_t)v[0]);
  r[1] = bf2f((uint16_t)v[1]);
  r[2] = bf2f((uint16_t)v[2]);
  r[3] = bf2f((uint16_t)v[3]);
  return r;
}
__device__ __forceinline__ bf16x4 gelu_bf4(f32x4 z) {
  f32x4 a;
  a[0] = mlp_gelu(z[0]);
  a[1] = mlp_gelu(z[1]);
  a[2] = mlp_gelu(z[2]);
  a[3] = mlp_gelu(z[3]);
  return pack_bf4(a);
}

struct MlpNoSink {
  __device__ __forceinline__ void operator()(int, int, f32x4, bf16x4) const {}
};

// Forward pass of one 16-sample tile on the whole wavefront. x0 / x1: this
// lane's B-operand inputs (features 4q..4q+3 and 16+4q..16+4q+3 of sample j).
// Returns the output block (lane (j, q): outputs 4q..4q+3 of sample j; outputs
// 4..15 are padding). sink(layer, block, z, a) receives each hidden block's
// pre-activation and BF16 activation. NB bounds the hidden blocks (h <= 16 NB).
template <int NB, class Sink>
__device__ __forceinline__ f32x4 mlp_tile_fwd(const MlpW& W, bf16x4 x0, bf16x4 x1, const Sink& sink) {
  const int lane = (int)(threadIdx.x & 63), j = lane & 15, q = lane >> 4;
  const int h = W.h, nb = h >> 4;
  bf16x4 hb[NB];
#pragma unroll
  for (int ob = 0; ob < NB; ob++) {
    if (ob < nb) {
      const uint16_t* wr = W.wt[0] + (int64_t)(16 * ob + j) * MLP_KIN + 4 * q;
      f32x4 acc = ld_f4(W.b[0] + 16 * ob + 4 * q);
      acc = mfma_bf16(ld_bf4(wr), x0, acc);
      acc = mfma_bf16(ld_bf4(wr + 16), x1, acc);
      hb[ob] = gelu_bf4(acc);
      sink(0, ob, acc, hb[ob]);
    }
  }
#pragma unroll
  for (int l = 1; l < 3; l++) {
    bf16x4 hn[NB];
#pragma unroll
    for (int ob = 0; ob < NB; ob++) {
      if (ob < nb) {
        const uint16_t* wr = W.wt[l] + (int64_t)(16 * ob + j) * h + 4 * q;
        f32x4 acc = ld_f4(W.b[l] + 16 * ob + 4 * q);
#pragma unroll
        for (int kb = 0; kb < NB; kb++)
          if (kb < nb) acc = mfma_bf16(ld_bf4(wr + 16 * kb), hb[kb], acc);
        hn[ob] = gelu_bf4(acc);
        sink(l, ob, acc, hn[ob]);
      }
    }
#pragma unroll
    for (int ob = 0; ob < NB; ob++) hb[ob] = hn[ob];
  }
  const uint16_t* wr = W.wt[3] + (int64_t)j * h + 4 * q;
  f32x4 acc = ld_f4(W.b[3] + 4 * q);
#pragma unroll
  for (int kb = 0; kb < NB; kb++)
    if (kb < nb) acc = mfma_bf16(ld_bf4(wr + 16 * kb), hb[kb], acc);
  return acc;
}

}  // namespace phk
