// Learned warm start: the step kernel instances that run the network, and the
// training kernels (forward and backward on the matrix cores, split-K weight
// gradients, Adam, the ring buffer of recorded contacts). See mlp.hpp for the
// network and mlp_dev.hpp for the MFMA layout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "kstep.hpp"
#include "mlp_host.hpp"
#include "spec_models.hpp"

using namespace phk;

namespace {

// ---------------------------------------------------------------- batch forward / backward
struct TileArgs {
  const float* feat;  // nsrc x MLP_NF
  const float* tgt;   // nsrc x MLP_NO, or null (forward only)
  int64_t nsrc, B;
  int identity;       // sample s reads row s; else a hashed row
  int backward;
  uint64_t seed;
  int64_t iter;
  uint16_t* x0;
  float* z;           // null: activations are not kept
  uint16_t* a;
  uint16_t* dz;
  float* y;
  float* loss;
};

// One wavefront per 16 samples: inputs, forward (keeping pre-activations and
// activations for the weight gradients), squared error, then the backward
// pass through the layers to dL/dZ of every hidden layer.
template <int NB>
__global__ void __launch_bounds__(64) k_mlp_tile(MlpW W, TileArgs T) {
  const int lane = (int)threadIdx.x, j = lane & 15, q = lane >> 4;
  const int h = W.h, nb = h >> 4;
  const int64_t s = (int64_t)blockIdx.x * 16 + j;
  const int64_t src =
      T.identity ? s : (int64_t)(mlp_mix(T.seed ^ mlp_mix(((uint64_t)T.iter << 32) ^ (uint64_t)s)) % (uint64_t)T.nsrc);
  const float* f = T.feat + src * MLP_NF;
  const bf16x4 zero = {0, 0, 0, 0};
  const bf16x4 x0 = pack_bf4(ld_f4(f + 4 * q));
  const bf16x4 x1 = q < 2 ? pack_bf4(ld_f4(f + 16 + 4 * q)) : zero;
  const int64_t BH = T.B * h;
  if (T.z) {
    st_bf4(T.x0 + s * MLP_KIN + 4 * q, x0);
    st_bf4(T.x0 + s * MLP_KIN + 16 + 4 * q, x1);
  }
  auto sink = [&](int l, int ob, f32x4 zz, bf16x4 aa) {
    if (T.z) {
      st_f4(T.z + l * BH + s * h + 16 * ob + 4 * q, zz);
      st_bf4(T.a + l * BH + s * h + 16 * ob + 4 * q, aa);
    }
  };
  const f32x4 y = mlp_tile_fwd<NB>(W, x0, x1, sink);
  if (T.y) st_f4(T.y + s * MLP_NOP + 4 * q, y);
  if (!T.tgt) return;
  f32x4 d = {0.0f, 0.0f, 0.0f, 0.0f};
  if (q == 0) {
    const f32x4 t = ld_f4(T.tgt + src * MLP_NO);
    d = y - t;
  }
  float sq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + d[3] * d[3];
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) sq += __shfl_xor(sq, o);
  if (lane == 0) T.loss[blockIdx.x] = sq;
  if (!T.backward) return;
  // L = sum (y - t)^2 / (4 B): dL/dy = 2 (y - t) / (4 B); padded outputs get 0
  const float sc = 2.0f / (4.0f * (float)T.B);
  const bf16x4 gy = pack_bf4(d * sc);
  st_bf4(T.dz + 3 * BH + s * MLP_NOP + 4 * q, gy);
  auto gelu_back = [&](f32x4 acc, int l, int ib) {
    const f32x4 zz = ld_f4(T.z + l * BH + s * h + 16 * ib + 4 * q);
    f32x4 r;
    r[0] = acc[0] * mlp_gelu_grad(zz[0]);
    r[1] = acc[1] * mlp_gelu_grad(zz[1]);
    r[2] = acc[2] * mlp_gelu_grad(zz[2]);
    r[3] = acc[3] * mlp_gelu_grad(zz[3]);
    const bf16x4 g = pack_bf4(r);
    st_bf4(T.dz + l * BH + s * h + 16 * ib + 4 * q, g);
    return g;
  };
  // through the output layer (W4: [h][16]) to dL/dZ3
  bf16x4 g[NB];
#pragma unroll
  for (int ib = 0; ib < NB; ib++) {
    if (ib < nb) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
      acc = mfma_bf16(ld_bf4(W.w[3] + (int64_t)(16 * ib + j) * MLP_NOP + 4 * q), gy, acc);
      g[ib] = gelu_back(acc, 2, ib);
    }
  }
  // through hidden layers 3 and 2 (W_l: [h][h]) to dL/dZ2, dL/dZ1
#pragma unroll
  for (int l = 2; l >= 1; l--) {
    bf16x4 gn[NB];
#pragma unroll
    for (int ib = 0; ib < NB; ib++) {
      if (ib < nb) {
        const uint16_t* wr = W.w[l] + (int64_t)(16 * ib + j) * h + 4 * q;
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ob = 0; ob < NB; ob++)
          if (ob < nb) acc = mfma_bf16(ld_bf4(wr + 16 * ob), g[ob], acc);
        gn[ib] = gelu_back(acc, l - 1, ib);
      }
    }
#pragma unroll
    for (int ib = 0; ib < NB; ib++) g[ib] = gn[ib];
  }
}

// dW_l = A_{l-1}^T dZ_l over one split of the batch, one 16 x 16 block per
// wavefront: A operand rows = inputs, k = samples; B operand k = samples,
// columns = outputs. Bias blocks (after the weight blocks) use an A operand
// whose row 0 is all ones, so that row 0 of the result is db_l = 1^T dZ_l.
struct WgArgs {
  const uint16_t* x0;
  const uint16_t* a;
  const uint16_t* dz;
  float* gp;
  int64_t B, np;
  int splits;
};
__global__ void __launch_bounds__(64) k_mlp_wgrad(int h, WgArgs A) {
  const int lane = (int)threadIdx.x, j = lane & 15, q = lane >> 4;
  const int nb = h >> 4;
  int blk = (int)blockIdx.x, l, ib, ob;
  bool bias = false;
  const int nwblk = 2 * nb + 2 * nb * nb + nb;
  if (blk >= nwblk) {  // bias blocks: layers 0..2 have nb each, layer 3 one
    bias = true;
    blk -= nwblk;
    l = blk < 3 * nb ? blk / nb : 3;
    ob = blk < 3 * nb ? blk % nb : 0;
    ib = 0;
  } else if (blk < 2 * nb) {
    l = 0;
    ib = blk / nb;
    ob = blk % nb;
  } else if (blk < 2 * nb + 2 * nb * nb) {
    blk -= 2 * nb;
    l = 1 + blk / (nb * nb);
    blk %= nb * nb;
    ib = blk / nb;
    ob = blk % nb;
  } else {
    l = 3;
    ib = blk - 2 * nb - 2 * nb * nb;
    ob = 0;
  }
  const int IN = mlp_in(l, h), OUT = mlp_out(l, h);
  const int64_t BH = A.B * h;
  const uint16_t* X = l == 0 ? A.x0 : A.a + (l - 1) * BH;
  const uint16_t* D = A.dz + l * BH;
  const int64_t chunk = A.B / A.splits, s0 = (int64_t)blockIdx.y * chunk;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int64_t s = s0; s < s0 + chunk; s += 16) {
    bf16x4 av, bv;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      av[v] = bias ? (short)(j == 0 ? 0x3f80 : 0) : (short)X[(s + 4 * q + v) * IN + 16 * ib + j];  // BF16 1.0
      bv[v] = (short)D[(s + 4 * q + v) * OUT + 16 * ob + j];
    }
    acc = mfma_bf16(av, bv, acc);
  }
  float* out = A.gp + (int64_t)blockIdx.y * A.np;
  if (bias) {
    if (q == 0) out[mlp_boff(l, h) + 16 * ob + j] = acc[0];  // row 0
    return;
  }
  out += mlp_woff(l, h);
#pragma unroll
  for (int v = 0; v < 4; v++) out[(int64_t)(16 * ib + 4 * q + v) * OUT + 16 * ob + j] = acc[v];
}

__global__ void __launch_bounds__(256) k_mlp_gsum(const float* gp, float* G, int64_t np, int splits) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= np) return;
  float s = 0.0f;
  for (int k = 0; k < splits; k++) s += gp[k * np + p];
  G[p] = s;
}

__global__ void __launch_bounds__(256) k_mlp_adam(float* P, float* m, float* v, const float* G, int64_t np, float lr,
                                                  float bc1, float bc2) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= np) return;
  const float g = G[p];
  const float mm = 0.9f * m[p] + 0.1f * g;
  const float vv = 0.999f * v[p] + 0.001f * (g * g);
  m[p] = mm;
  v[p] = vv;
  P[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + 1e-8f);
}

__global__ void __launch_bounds__(256) k_mlp_tobf16(int h, const float* P, uint16_t* wbf, uint16_t* wtbf) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  for (int l = 0; l < 4; l++) {
    const int64_t w0 = mlp_woff(l, h), IN = mlp_in(l, h), OUT = mlp_out(l, h);
    if (p >= w0 && p < w0 + IN * OUT) {
      const int64_t i = (p - w0) / OUT, o = (p - w0) % OUT;
      const uint16_t b = f2bf(P[p]);
      wbf[mlp_wbf_off(l, h) + i * OUT + o] = b;
      wtbf[mlp_wbf_off(l, h) + o * IN + i] = b;
    }
  }
}

// Ring append in env order: one workgroup scans the per-env counts into
// offsets, then one thread per (env, contact) copies its record.
__global__ void __launch_bounds__(1024) k_mlp_scan(const int* cnt, int64_t begin, int64_t end, int64_t* off,
                                                   int64_t* total) {
  __shared__ int sc[1024];
  const int tid = (int)threadIdx.x;
  int64_t run = 0;
  for (int64_t base = begin; base < end; base += 1024) {
    const int64_t e = base + tid;
    const int c = e < end ? cnt[e] : 0;
    sc[tid] = c;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const int add = tid >= o ? sc[tid - o] : 0;
      __syncthreads();
      sc[tid] += add;
      __syncthreads();
    }
    if (e < end) off[e - begin] = run + sc[tid] - c;
    run += sc[1023];
    __syncthreads();
  }
  if (tid == 0) *total = run;
}

__global__ void __launch_bounds__(256) k_mlp_scatter(const int* cnt, const float* feat, const float* tgt,
                                                     int64_t begin, int64_t end, const int64_t* off, float* rfeat,
                                                     float* rtgt, int64_t cap, int64_t head) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t e = begin + t / MLP_CL;
  const int k = (int)(t % MLP_CL);
  if (e >= end || k >= cnt[e]) return;
  const int64_t dst = (head + off[e - begin] + k) % cap;
  const int64_t srow = e * MLP_CL + k;
  for (int i = 0; i < MLP_NF; i++) rfeat[dst * MLP_NF + i] = feat[srow * MLP_NF + i];
  for (int i = 0; i < MLP_NO; i++) rtgt[dst * MLP_NO + i] = tgt[srow * MLP_NO + i];
}

int pick_splits(int64_t B, int want) {
  int s = (int)std::max<int64_t>(1, std::min<int64_t>(want, B / 1024));
  while (s > 1 && (B / 16) % s != 0) s--;
  return s;
}

}  // namespace

namespace phm {

const void* step_kernel(int K, int spec) {
  if (spec == 1 && K == 4) return (const void*)k_step_mlp<4, SpecAnt>;
  return K == 4 ? (const void*)k_step_mlp<4, DModel>
                : K == 2 ? (const void*)k_step_mlp<2, DModel> : (const void*)k_step_mlp<1, DModel>;
}

void launch_step(int K, int spec, unsigned grid, size_t lds, hipStream_t st, const DModel* M, const DState* S,
                 int64_t begin, int64_t end, int64_t nstep, const double* ctrl, double* qpos_out) {
  if (spec == 1 && K == 4)
    hipLaunchKernelGGL((k_step_mlp<4, SpecAnt>), dim3(grid), dim3(64), lds, st, M, S, begin, end, nstep, ctrl,
                       qpos_out);
  else if (K == 4)
    hipLaunchKernelGGL((k_step_mlp<4, DModel>), dim3(grid), dim3(64), lds, st, M, S, begin, end, nstep, ctrl,
                       qpos_out);
  else if (K == 2)
    hipLaunchKernelGGL((k_step_mlp<2, DModel>), dim3(grid), dim3(64), lds, st, M, S, begin, end, nstep, ctrl,
                       qpos_out);
  else
    hipLaunchKernelGGL((k_step_mlp<1, DModel>), dim3(grid), dim3(64), lds, st, M, S, begin, end, nstep, ctrl,
                       qpos_out);
}

MlpW view(const Net& n, int mode) {
  MlpW W{};
  W.mode = mode;
  W.h = n.h;
  for (int l = 0; l < 4; l++) {
    W.wt[l] = n.wtbf + mlp_wbf_off(l, n.h);
    W.w[l] = n.wbf + mlp_wbf_off(l, n.h);
    W.b[l] = n.P + mlp_boff(l, n.h);
  }
  return W;
}

void to_bf16(const Net& n, hipStream_t st) {
  const int64_t np = mlp_nparam(n.h);
  hipLaunchKernelGGL(k_mlp_tobf16, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, n.h, n.P, n.wbf, n.wtbf);
}

void forward_backward(const Net& n, const Batch& b, const float* feat, const float* tgt, int64_t nsrc, int64_t B,
                      bool identity, uint64_t seed, int64_t iter, bool backward, hipStream_t st) {
  TileArgs T{feat, tgt, nsrc, B, identity ? 1 : 0, backward ? 1 : 0, seed, iter, b.x0, backward ? b.z : nullptr,
             b.a, b.dz, b.y, b.loss};
  hipLaunchKernelGGL(k_mlp_tile<MLP_HMAX / 16>, dim3((unsigned)(B / 16)), dim3(64), 0, st, view(n, 0), T);
  if (!backward) return;
  const int64_t np = mlp_nparam(n.h);
  const int splits = pick_splits(B, b.splits);
  const int nb = n.h / 16;
  WgArgs A{b.x0, b.a, b.dz, b.gp, B, np, splits};
  hipLaunchKernelGGL(k_mlp_wgrad, dim3((unsigned)(2 * nb + 2 * nb * nb + nb + 3 * nb + 1), (unsigned)splits), dim3(64),
                     0, st, n.h, A);
  hipLaunchKernelGGL(k_mlp_gsum, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, b.gp, n.G, np, splits);
}

void adam(const Net& n, int64_t t, float lr, hipStream_t st) {
  const int64_t np = mlp_nparam(n.h);
  const float bc1 = (float)(1.0 - std::pow(0.9, (double)t));
  const float bc2 = (float)(1.0 - std::pow(0.999, (double)t));
  hipLaunchKernelGGL(k_mlp_adam, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, n.P, n.m, n.v, n.G, np, lr,
                     bc1, bc2);
  to_bf16(n, st);
}

void collect(const int* cnt, const float* feat, const float* tgt, int64_t begin, int64_t end, float* rfeat,
             float* rtgt, int64_t cap, int64_t head, int64_t* d_scratch, hipStream_t st) {
  // d_scratch: the total, then one offset per env
  hipLaunchKernelGGL(k_mlp_scan, dim3(1), dim3(1024), 0, st, cnt, begin, end, d_scratch + 1, d_scratch);
  const int64_t threads = (end - begin) * MLP_CL;
  hipLaunchKernelGGL(k_mlp_scatter, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, cnt, feat, tgt, begin,
                     end, d_scratch + 1, rfeat, rtgt, cap, head);
}

}  // namespace phm
