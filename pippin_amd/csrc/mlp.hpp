// Learned warm start for the Newton solver. It has no counterpart in the
// reference, whose warm start is the previous acceleration, accepted when its
// cost is lower than that of the unconstrained one (newton.rs:238-245).
//
// A small MLP maps 24 features of a touching contact to the four pyramid-row
// forces of that contact, as y_r = log1p(lambda_r / D_r) (D_r the row's
// impedance weight, so lambda_r / D_r = max(0, aref_r - J_r a) at the solution
// a). Inside the step kernel the predicted forces of an env's first MLP_CL
// touching contacts give a third warm-start candidate
//     a_mlp = M^-1 (smooth + sum_r J_r^T lambda_r) = a0 + M^-1 J^T lambda,
// which Newton takes only if its cost is lower than the cost of the
// reference's choice; the iteration and its stopping rule are unchanged. With
// logging on, the step kernel also records each contact's features and its
// solved forces for online training.
//
// Network: 24 -> H -> H -> H -> 4, GELU (tanh form) after the hidden layers,
// BF16 weights and activations, FP32 accumulation (v_mfma_f32_16x16x16_bf16).
// Inputs are padded to 32 and outputs to 16 with zero weights.
#pragma once
#include <cstdint>
#include <cstring>

#include "dmath.hpp"

namespace phk {

constexpr int MLP_NF = 24;          // features per contact
constexpr int MLP_KIN = 32;         // features padded to two 16-wide k-blocks
constexpr int MLP_NO = 4;           // outputs: the contact's four pyramid rows
constexpr int MLP_NOP = 16;         // outputs padded to one 16-row block
constexpr int MLP_CL = 16;          // contacts per env with features and predictions
constexpr int MLP_HMAX = 256;       // widest hidden layer (training and batch inference)
constexpr int MLP_HMAX_STEP = 64;   // widest hidden layer inside the step kernel
enum { MLP_LOG = 1, MLP_WARM = 2 };

// Device view of the network, inside DModel. Layer l (0..3) has fan-in
// IN_l = 32, H, H, H and fan-out OUT_l = H, H, H, 16.
struct MlpW {
  int mode;                 // MLP_LOG | MLP_WARM, 0 = off
  int h;                    // hidden width, a multiple of 16
  const uint16_t* wt[4];    // BF16, [OUT_l][IN_l] (forward operand)
  const uint16_t* w[4];     // BF16, [IN_l][OUT_l] (backward operand)
  const float* b[4];        // FP32 biases, OUT_l
};

PHD int mlp_in(int l, int h) { return l == 0 ? MLP_KIN : h; }
PHD int mlp_out(int l, int h) { return l == 3 ? MLP_NOP : h; }
// Offsets of W_l ([IN_l][OUT_l]) and b_l in the flat FP32 parameter vector
// W1 b1 W2 b2 W3 b3 W4 b4.
PHD int64_t mlp_woff(int l, int h) {
  int64_t o = 0;
  for (int k = 0; k < l; k++) o += (int64_t)mlp_in(k, h) * mlp_out(k, h) + mlp_out(k, h);
  return o;
}
PHD int64_t mlp_boff(int l, int h) { return mlp_woff(l, h) + (int64_t)mlp_in(l, h) * mlp_out(l, h); }
PHD int64_t mlp_nparam(int h) { return mlp_woff(4, h); }
// Offset of W_l inside the packed BF16 weight arrays (W_l, or its transpose).
PHD int64_t mlp_wbf_off(int l, int h) {
  int64_t o = 0;
  for (int k = 0; k < l; k++) o += (int64_t)mlp_in(k, h) * mlp_out(k, h);
  return o;
}
PHD int64_t mlp_nweight(int h) { return mlp_wbf_off(4, h); }

// FP32 -> BF16, round to nearest even (NaN stays NaN).
PHD uint16_t f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
PHD float bf2f(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// GELU, tanh form: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))), and its derivative.
PHD float mlp_gelu(float x) {
  float u = 0.7978845608028654f * (x + 0.044715f * (x * x * x));
  return 0.5f * x * (1.0f + tanhf(u));
}
PHD float mlp_gelu_grad(float x) {
  float u = 0.7978845608028654f * (x + 0.044715f * (x * x * x));
  float t = tanhf(u);
  float du = 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * (x * x));
  return 0.5f * (1.0f + t) + 0.5f * x * ((1.0f - t * t) * du);
}

// Signed log compression of features: sign(x) log(1 + |x|).
PHD Real mlp_slog(Real x) { return x < 0.0 ? -log1p(-x) : log1p(x); }

// 64-bit mix (SplitMix64 finaliser), for batch sampling.
PHD uint64_t mlp_mix(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

}  // namespace phk
