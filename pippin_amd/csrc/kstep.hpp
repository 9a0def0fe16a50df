// The step kernels: the execution policy for K envs per wavefront, the plain
// step kernel and its variant with the learned warm start.
#pragma once
#include <hip/hip_runtime.h>

#include "kbody.hpp"
#include "mlp_dev.hpp"

namespace phk {

// K environments per workgroup of one 64-lane wavefront: env g of the
// workgroup owns lanes [g * 64/K, (g + 1) * 64/K) and its own slice of the LDS.
// Loops over an env's items stride over its lanes; the dense linear algebra
// uses the first 16 of them (one DPP row).
template <int K>
struct DevXK : DevLanes {
  static constexpr int G = 64 / K;
  template <class F>
  __device__ __forceinline__ static void par(int n, F&& f) {
    for (int i = (int)(threadIdx.x % G); i < n; i += G) f(i);
    __syncthreads();
  }
  __device__ __forceinline__ static void sync() { __syncthreads(); }
};
using DevX = DevXK<1>;

// The policy of the learned-warm-start kernel: DevXK plus the network over
// the wavefront's recorded contacts. Tile t holds contacts t * 16/K ..
// (t + 1) * 16/K - 1 of each of the K envs, one sample column per contact;
// tiles run while any env of the wavefront has contacts left.
template <int K>
struct DevXKM : DevXK<K> {
  static constexpr int G = 64 / K;
  template <class MT>
  __device__ static void mlp_tiles(const MT& M, const DState& S, int64_t e_self, int cnt) {
    constexpr int SPE = 16 / K;
    if (M.mlp.h > MLP_HMAX_STEP) {  // wider networks are trained and run in batches only
      __syncthreads();
      return;
    }
    const int lane = (int)threadIdx.x, g = lane / G, j = lane & 15, q = lane >> 4;
    const int64_t e0 = e_self - g;
    int mx = cnt;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) mx = max(mx, __shfl_xor(mx, o));
    const int gs = j / SPE;  // env of this lane's sample column
    const int cs = __shfl(cnt, gs * G);
    const int nt = (mx + SPE - 1) / SPE;
    const bf16x4 zero = {0, 0, 0, 0};
    for (int t = 0; t < nt; t++) {
      const int c = t * SPE + j % SPE;
      const bool have = c < cs;
      const int64_t row = (e0 + gs) * MLP_CL + c;
      bf16x4 x0 = zero, x1 = zero;
      if (have) {
        const float* f = S.mlp_feat.p + row * MLP_NF;
        x0 = pack_bf4(ld_f4(f + 4 * q));
        if (q < 2) x1 = pack_bf4(ld_f4(f + 16 + 4 * q));
      }
      const f32x4 y = mlp_tile_fwd<MLP_HMAX_STEP / 16>(M.mlp, x0, x1, MlpNoSink{});
      if (q == 0 && have) st_f4(S.mlp_pred.p + row * MLP_NO, y);
    }
    __syncthreads();
  }
};

// Advance envs [begin, end) by nstep steps, K per workgroup. With ctrl_seq
// (envs x nstep x nu) the controls change every step and qpos_out
// (envs x nstep x nq) receives qpos after each step (Batch::rollout, batch.rs:46-59).
// MT is DModel (sizes and layout read at run time) or a model-specialised
// struct derived from it whose sizes and LDS layout are compile-time constants
// (spec_models.hpp, generated at build time); both read the same buffers.
template <int K, class MT>
__global__ void __launch_bounds__(64) k_step(const DModel* __restrict__ Mp, const DState* __restrict__ Sp,
                                             int64_t begin, int64_t end, int64_t nstep, const Real* ctrl_seq,
                                             Real* qpos_out) {
  const MT& M = *static_cast<const MT*>(Mp);
  const DState& S = *Sp;
  extern __shared__ double lds[];
  const int g = (int)threadIdx.x / DevXK<K>::G;
  const int64_t rel = (int64_t)blockIdx.x * K + g;
  if (begin + rel >= end) return;  // tail of the last workgroup
  Env E;
  E.s = lds + (int64_t)g * (M.L.bytes / 8);
  E.si = reinterpret_cast<int*>(E.s + M.L.ndouble);
  E.e = begin + rel;
  E.var = M.env_var[E.e];
  run_env<DevXK<K>>(M, S, E, nstep, ctrl_seq, qpos_out, rel);
}

// Persistent work-stealing variant of k_step: the grid is what the device
// holds at once, and each workgroup takes the next slot of K envs from
// *counter until the slots run out. Slot s holds envs order[s*K .. s*K+K-1]
// (offsets from begin; the identity without order). k_bucket writes an order
// that puts envs of similar cost in one wavefront, costliest first, so that a
// wavefront waits less for its slowest env and the long slots start early.
// Every env's arithmetic is the same under any order. ctrl_seq / qpos_out as
// for k_step (rows indexed by the env's offset from begin).
template <int K, class MT>
__global__ void __launch_bounds__(64) k_step_ws(const DModel* __restrict__ Mp, const DState* __restrict__ Sp,
                                                int64_t begin, int64_t end, int64_t nstep, const Real* ctrl_seq,
                                                Real* qpos_out, const int* __restrict__ order, unsigned* counter) {
  const MT& M = *static_cast<const MT*>(Mp);
  const DState& S = *Sp;
  extern __shared__ double lds[];
  const int g = (int)threadIdx.x / DevXK<K>::G;
  const int64_t n = end - begin;
  const int64_t nslots = (n + K - 1) / K;
  for (;;) {
    unsigned slot = 0;
    if (threadIdx.x == 0) slot = atomicAdd(counter, 1u);
    slot = __shfl(slot, 0);
    if ((int64_t)slot >= nslots) break;
    const int64_t idx = (int64_t)slot * K + g;
    if (idx < n) {
      const int64_t rel = order ? (int64_t)order[idx] : idx;
      Env E;
      E.s = lds + (int64_t)g * (M.L.bytes / 8);
      E.si = reinterpret_cast<int*>(E.s + M.L.ndouble);
      E.e = begin + rel;
      E.var = M.env_var[E.e];
      run_env<DevXK<K>>(M, S, E, nstep, ctrl_seq, qpos_out, rel);
    }
    __syncthreads();
  }
}

// The PGS step split around its sweeps (run_pgs_front, run_pgs_back in
// kbody.hpp), one env per workgroup of one wavefront, for physics step t of
// nstep (K = 1): the front and back keep the step kernel's LDS working set; the sweeps
// kernel needs no LDS and few registers, so many of its wavefronts share a SIMD
// and hide each other's dependent chains. vel and vpos hold nv doubles per env
// (indexed by env id).
template <int K>
__global__ void __launch_bounds__(64) k_pgs_front(const DModel* __restrict__ Mp, const DState* __restrict__ Sp,
                                                  int64_t begin, int64_t end, int64_t nstep, int64_t t,
                                                  const Real* ctrl_seq, Real* vel, Real* vpos) {
  const DModel& M = *Mp;
  const DState& S = *Sp;
  extern __shared__ double lds[];
  const int64_t rel = (int64_t)blockIdx.x;
  if (begin + rel >= end) return;
  Env E;
  E.s = lds;
  E.si = reinterpret_cast<int*>(E.s + M.L.ndouble);
  E.e = begin + rel;
  E.var = M.env_var[E.e];
  static_assert(K == 1, "one env per workgroup");
  run_pgs_front<DevXK<1>>(M, S, E, nstep, t, ctrl_seq, rel, vel + E.e * M.nv, vpos + E.e * M.nv);
}

template <int K>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) k_pgs_sweeps(const DModel* __restrict__ Mp, const DState* __restrict__ Sp,
                                                   int64_t begin, int64_t end, Real* vel, Real* vpos) {
#if defined(__HIP_DEVICE_COMPILE__)
  const DModel& M = *Mp;
  const DState& S = *Sp;
  const int64_t e = begin + (int64_t)blockIdx.x;
  if (e >= end || S.status[e] != 0) return;
  const int nr = S.nrows[e];
  if (nr <= 0) return;
  const int nv = M.nv;
  Real* J = (S.gjac + e * (int64_t)M.nrmax * nv).p;
  Real* mj = (S.gmj + e * (int64_t)M.nrmax * nv).p;
  Real* rw = (S.grow + e * (int64_t)M.nrmax * 10).p;
  Real* v = vel + e * nv;
  Real* vp = vpos + e * nv;
  const PgsConsts C = pgs_consts(M);
  if (nr <= PGS_WAVE_ROWS) {
    const int lane = (int)threadIdx.x;
    const LaneVel r = pgs_sweeps_body(C, J, mj, rw, nr, nv, lane < nv ? v[lane] : 0.0);
    if (lane < nv) {
      v[lane] = r.v;
      vp[lane] = r.vpos;
    }
  } else if (threadIdx.x == 0) {
    pgs_sweeps_serial(C, J, mj, rw, nr, nv, v, vp);
  }
#endif
}

template <int K>
__global__ void __launch_bounds__(64) k_pgs_back(const DModel* __restrict__ Mp, const DState* __restrict__ Sp,
                                                 int64_t begin, int64_t end, int64_t nstep, int64_t t, Real* qpos_out,
                                                 const Real* vel, const Real* vpos) {
  const DModel& M = *Mp;
  const DState& S = *Sp;
  extern __shared__ double lds[];
  const int64_t rel = (int64_t)blockIdx.x;
  if (begin + rel >= end) return;
  Env E;
  E.s = lds;
  E.si = reinterpret_cast<int*>(E.s + M.L.ndouble);
  E.e = begin + rel;
  E.var = M.env_var[E.e];
  static_assert(K == 1, "one env per workgroup");
  run_pgs_back<DevXK<1>>(M, S, E, nstep, t, qpos_out, rel, vel + E.e * M.nv, vpos + E.e * M.nv);
}

// Contact-count bucketing for k_step_ws, one workgroup: envs [begin, end)
// keyed by the Newton iterations of their last launch (summed over its steps)
// and their last contact count, grouped by key, costliest key first, into
// order (offsets from begin). Also resets the work-stealing counter for the
// launch that follows on the same stream.
constexpr int BUCKETS = 64;
static __global__ void __launch_bounds__(1024) k_bucket(const DState* __restrict__ Sp, int64_t begin, int64_t end,
                                                 int* __restrict__ order, unsigned* counter) {
  const DState& S = *Sp;
  __shared__ unsigned hist[BUCKETS];
  __shared__ unsigned off[BUCKETS];
  const int t = (int)threadIdx.x;
  const int64_t n = end - begin;
  if (t < BUCKETS) hist[t] = 0;
  __syncthreads();
  auto key = [&](int64_t e) {
    // iterations dominate an env's cost; contacts break ties within them
    int it = S.cost[e], nc = S.ncon[e];
    it = it < 0 ? 0 : (it > 31 ? 31 : it);
    nc = nc < 0 ? 0 : (nc > 1 ? 1 : nc);
    return it * 2 + nc;
  };
  for (int64_t i = t; i < n; i += blockDim.x) atomicAdd(&hist[key(begin + i)], 1u);
  __syncthreads();
  if (t == 0) {
    unsigned o = 0;
    for (int b = BUCKETS - 1; b >= 0; b--) {
      off[b] = o;
      o += hist[b];
    }
    counter[0] = 0;
  }
  __syncthreads();
  for (int64_t i = t; i < n; i += blockDim.x) {
    unsigned p = atomicAdd(&off[key(begin + i)], 1u);
    order[p] = (int)i;
  }
}

// Resets the work-stealing counter when no bucketing pass runs before k_step_ws
// (a kernel, not a memset, so that the launch sequence can be captured in a graph).
static __global__ void k_reset_counter(unsigned* counter) { counter[0] = 0; }

// The step with the learned warm start (mlp.hpp). Lane groups past the end
// stay in the loop without an env, since the network needs the whole wavefront.
template <int K, class MT>
__global__ void __launch_bounds__(64) k_step_mlp(const DModel* __restrict__ Mp, const DState* __restrict__ Sp,
                                                 int64_t begin, int64_t end, int64_t nstep, const Real* ctrl_seq,
                                                 Real* qpos_out) {
  const MT& M = *static_cast<const MT*>(Mp);
  const DState& S = *Sp;
  extern __shared__ double lds[];
  const int g = (int)threadIdx.x / DevXK<K>::G;
  const int64_t rel = (int64_t)blockIdx.x * K + g;
  const bool live = begin + rel < end;
  Env E;
  E.s = lds + (int64_t)g * (M.L.bytes / 8);
  E.si = reinterpret_cast<int*>(E.s + M.L.ndouble);
  E.e = begin + rel;
  E.var = live ? M.env_var[E.e] : 0;
  run_env_mlp<DevXKM<K>>(M, S, E, nstep, ctrl_seq, qpos_out, rel, live);
}

}  // namespace phk
