// ms_reduce.hpp -- deterministic two-stage reductions (device code).
//
// Stage 1 (inside the producing kernel): wave64 shuffles, then LDS across
// the 4 waves of a workgroup; thread 0 writes the workgroup's partials.
// Stage 2 (k_finish, one workgroup of 1024 threads): sums the partials in a
// fixed order and applies the finisher functor (scalar recurrences of the
// solver) on one thread.  The kernel boundary between the two stages provides the memory
// ordering, so no device-scope fence or atomic is needed: on a multi-XCD
// CDNA device an agent-scope release fence in every workgroup would write
// back the XCD's L2 each time.  Results depend only on the grid size.
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

// Workgroup reduction of NV values per thread (BS threads): wave64 shuffles,
// then the wave results in LDS, combined by thread 0 in wave order.  The
// result is valid on thread 0.
template <class Op, int NV, class T, int BS = kBlock>
__device__ __forceinline__ void block_reduce(T (&v)[NV], T *lds) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int j = 0; j < NV; j++) v[j] = Op::apply(v[j], __shfl_down(v[j], o, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = BS >> 6;
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

// Stage 1: one partial per workgroup (NV values at partial[idx*NV], idx =
// blockIdx by default).
template <class Op, int NV, class T>
__device__ __forceinline__ void block_partials(T (&v)[NV], T *partial, int idx = -1) {
  __shared__ T lds[NV * (kBlock / 64)];
  block_reduce<Op, NV, T>(v, lds);
  if (idx < 0) idx = blockIdx.x;
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < NV; j++) partial[idx * NV + j] = v[j];
}

// Stage 2 runs in one workgroup of kFinThreads threads.  The partials were
// written by workgroups on every XCD, so each load is a far round trip:
// every thread issues its loads four at a time before combining them (in
// index order), which keeps the reduction deterministic and hides latency.
constexpr int kFinThreads = 1024;

template <class Op, int NV, class T>
__device__ __forceinline__ void fin_gather(int nparts, const T *partial, T (&acc)[NV]) {
#pragma unroll
  for (int j = 0; j < NV; j++) acc[j] = Op::template init<T>();
  int b = threadIdx.x;
  for (; b + 3 * kFinThreads < nparts; b += 4 * kFinThreads) {
    T x[4][NV];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int j = 0; j < NV; j++) x[u][j] = partial[(b + u * kFinThreads) * NV + j];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int j = 0; j < NV; j++) acc[j] = Op::apply(acc[j], x[u][j]);
  }
  for (; b < nparts; b += kFinThreads)
#pragma unroll
    for (int j = 0; j < NV; j++) acc[j] = Op::apply(acc[j], partial[b * NV + j]);
}

template <class Op, int NV, class T, class Fin>
__global__ __launch_bounds__(kFinThreads) void k_finish(int nparts, const T *partial, Fin fin) {
  if (fin.skip()) return;
  T acc[NV];
  fin_gather<Op, NV, T>(nparts, partial, acc);
  __shared__ T lds[NV * (kFinThreads / 64)];
  block_reduce<Op, NV, T, kFinThreads>(acc, lds);
  if (threadIdx.x == 0) fin(acc);
}

__device__ __forceinline__ double tabs(double x) { return fabs(x); }
__device__ __forceinline__ float tabs(float x) { return fabsf(x); }
__device__ __forceinline__ double tsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float tsqrt(float x) { return sqrtf(x); }

// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx % 8).
// xcd_block maps them so that each XCD works on one contiguous range of
// logical blocks; the neighbour gathers of its rows then mostly hit its own
// L2 instead of being fetched again by every XCD.  A bijection of [0, nb).
inline constexpr int kXcd = 8;
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int q = nb / kXcd, r = nb % kXcd, x = b % kXcd;
  return x * q + min(x, r) + b / kXcd;
}

}  // namespace ms
