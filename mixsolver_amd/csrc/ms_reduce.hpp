// ms_reduce.hpp -- deterministic two-stage workgroup reductions (device code).
//
// Stage 1: wave64 shuffles, then LDS across the 4 waves of a workgroup.
// Stage 2: every workgroup publishes its partials; the workgroup that takes
// the last ticket sums all partials in a fixed order and runs the finisher on
// one thread.  Results depend only on the grid size, never on scheduling.
#pragma once
#include "ms_common.hpp"

namespace ms {

struct OpSum {
  template <class T> __device__ static T apply(T a, T b) { return a + b; }
  template <class T> __device__ static T init() { return T(0); }
};
struct OpMax {
  template <class T> __device__ static T apply(T a, T b) { return a > b ? a : b; }
  template <class T> __device__ static T init() { return T(0); }
};

template <class Op, int NV, class T>
__device__ __forceinline__ void block_reduce(T (&v)[NV], T *lds) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int j = 0; j < NV; j++) v[j] = Op::apply(v[j], __shfl_down(v[j], o, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = kBlock >> 6;
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < NV; j++) lds[w * NV + j] = v[j];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < NV; j++) {
      T s = lds[j];
      for (int q = 1; q < nw; q++) s = Op::apply(s, lds[q * NV + j]);
      v[j] = s;
    }
  __syncthreads();
}

template <class Op, int NV, class T, class Fin>
__device__ __forceinline__ void grid_reduce(T (&v)[NV], T *partial, unsigned *ticket, Fin fin) {
  __shared__ T lds[NV * (kBlock / 64)];
  __shared__ int last;
  block_reduce<Op, NV, T>(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; j++) partial[blockIdx.x * NV + j] = v[j];
    __threadfence();
    unsigned t = atomicAdd(ticket, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  T acc[NV];
#pragma unroll
  for (int j = 0; j < NV; j++) acc[j] = Op::template init<T>();
  const volatile T *vp = partial;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += kBlock)
#pragma unroll
    for (int j = 0; j < NV; j++) acc[j] = Op::apply(acc[j], (T)vp[b * NV + j]);
  block_reduce<Op, NV, T>(acc, lds);
  if (threadIdx.x == 0) {
    fin(acc);
    *ticket = 0u;
    __threadfence();
  }
}

}  // namespace ms

namespace ms {
__device__ __forceinline__ double tabs(double x) { return fabs(x); }
__device__ __forceinline__ float tabs(float x) { return fabsf(x); }
__device__ __forceinline__ double tsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float tsqrt(float x) { return sqrtf(x); }
}  // namespace ms
