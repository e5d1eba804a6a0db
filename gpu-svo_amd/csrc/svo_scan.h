// svo_scan.h -- wave64 / workgroup scan helpers (256-thread blocks).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace svo {

template <typename T>
__device__ inline T wave_incl_scan(T v) {
  const uint32_t lane = __lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const T o = __shfl_up(v, off);
    if (lane >= (uint32_t)off) v += o;
  }
  return v;
}

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Exclusive scan over a 256-thread block; `total` receives the block sum.
// `tmp` must hold 4 values.  Contains __syncthreads().
template <typename T>
__device__ inline T block_excl_scan256(T v, T* tmp, T& total) {
  const uint32_t w = threadIdx.x >> 6;
  const T inc = wave_incl_scan(v);
  if ((threadIdx.x & 63) == 63) tmp[w] = inc;
  __syncthreads();
  T pre = 0, tot = 0;
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    const T t = tmp[k];
    if (k < w) pre += t;
    tot += t;
  }
  __syncthreads();
  total = tot;
  return pre + inc - v;
}

// Three-component exclusive scan (vertex, edge, face counts).
struct U64x3 {
  uint64_t a, b, c;
};

}  // namespace svo
