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

// Inclusive scan of a 32-bit value over the wave through DPP lane moves (no
// LDS traffic): shifts by 1, 2, 4 and 8 lanes inside each row of 16, then
// lane 15 of each row added into the next row, then lane 31 into rows 2-3.
// Lanes that a move does not reach receive 0.
__device__ inline uint32_t wave_incl_scan_dpp(uint32_t v) {
  int x = (int)v;
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, true);  // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, true);  // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, true);  // row_shr:4
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, true);  // row_shr:8
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false);  // row_bcast:15, rows 1, 3
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, false);  // row_bcast:31, rows 2, 3
  return (uint32_t)x;
}

// Exclusive scan of two 32-bit values over a 256-thread block, in place;
// ta / tb receive the block sums.  `tmp` holds 8 values.  One barrier: a
// caller that scans again with the same `tmp` must put a barrier between the
// calls.
__device__ inline void block_excl_scan256_2(uint32_t& a, uint32_t& b, uint32_t* tmp,
                                            uint32_t& ta, uint32_t& tb) {
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t ia = wave_incl_scan_dpp(a), ib = wave_incl_scan_dpp(b);
  if ((threadIdx.x & 63) == 63) {
    tmp[w] = ia;
    tmp[4 + w] = ib;
  }
  __syncthreads();
  uint32_t pa = 0, pb = 0, sa = 0, sb = 0;
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    const uint32_t xa = tmp[k], xb = tmp[4 + k];
    if (k < w) {
      pa += xa;
      pb += xb;
    }
    sa += xa;
    sb += xb;
  }
  ta = sa;
  tb = sb;
  a = pa + ia - a;
  b = pb + ib - b;
}

// Three-component exclusive scan (vertex, edge, face counts).
struct U64x3 {
  uint64_t a, b, c;
};

// Loads cnt <= N consecutive keys p[0 .. cnt) into kk (the rest zero); with
// vec set (p 16-byte aligned) a full run is loaded as 16-byte vectors.
template <typename K, int N>
__device__ inline void load_consecutive(const K* p, bool vec, uint32_t cnt, K (&kk)[N]) {
  static_assert((N * sizeof(K)) % 16 == 0, "runs are whole 16-byte vectors");
  if (vec && cnt == (uint32_t)N) {
    const uint4* q = (const uint4*)p;
    constexpr int PER = 16 / sizeof(K);
#pragma unroll
    for (int v = 0; v < N / PER; ++v) {
      const uint4 x = q[v];
      const K* e = (const K*)&x;
#pragma unroll
      for (int j = 0; j < PER; ++j) kk[v * PER + j] = e[j];
    }
  } else {
#pragma unroll
    for (int j = 0; j < N; ++j) kk[j] = (uint32_t)j < cnt ? p[j] : (K)0;
  }
}

// ------------------------------------------------------ decoupled look-back
// Single-pass scans over tiles that workgroups claim in order with an atomic
// ticket.  Tile t publishes its count in a 64-bit status word -- bits 62-63
// flag (1 = the tile's own count, 2 = inclusive prefix), bits 32-61 epoch,
// bits 0-31 value -- and sums the words of earlier tiles, nearest first,
// until it meets an inclusive one.  Words whose epoch differs from the
// caller's have not been written by this pass: the words of the tiles in use
// are zeroed once and every pass over them uses its own nonzero epoch.
// Every earlier tile must be held by a running workgroup that publishes its
// own count without waiting on a later tile; then every wait ends.
constexpr uint64_t LB_AGG = 1ull << 62, LB_INCL = 2ull << 62, LB_FLAGS = 3ull << 62;
constexpr int LB_WINDOW = 16;  // words loaded per step: one latency covers them

__device__ inline uint64_t lb_word(uint64_t flag, uint32_t epoch, uint32_t v) {
  return flag | ((uint64_t)(epoch & 0x3FFFFFFFu) << 32) | v;
}

// Publishes `count` of tile `tile` at words[tile * stride] and returns the
// sum of the counts of tiles 0 .. tile-1.
__device__ inline uint32_t lookback_exclusive(uint64_t* words, uint64_t stride, uint64_t tile,
                                              uint32_t count, uint32_t epoch) {
  epoch &= 0x3FFFFFFFu;
  uint64_t* own = words + tile * stride;
  if (tile == 0) {
    __hip_atomic_store(own, lb_word(LB_INCL, epoch, count), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    return 0;
  }
  __hip_atomic_store(own, lb_word(LB_AGG, epoch, count), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  uint32_t prefix = 0;
  int64_t t = (int64_t)tile - 1;
  while (true) {
    uint64_t wv[LB_WINDOW];
#pragma unroll
    for (int j = 0; j < LB_WINDOW; ++j)
      wv[j] = t - j >= 0 ? __hip_atomic_load(words + (uint64_t)(t - j) * stride,
                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                         : 0ull;
    int used = 0;
    bool stop = false, found = false;
#pragma unroll
    for (int j = 0; j < LB_WINDOW; ++j) {
      if (!stop) {
        const uint64_t x = wv[j];
        if ((x & LB_FLAGS) == 0 || (uint32_t)((x >> 32) & 0x3FFFFFFFu) != epoch) {
          stop = true;  // not yet published: load it again
        } else {
          prefix += (uint32_t)x;
          ++used;
          if ((x & LB_FLAGS) == LB_INCL) stop = found = true;
        }
      }
    }
    if (found) break;
    t -= used;  // tile 0 publishes an inclusive word, so t stays >= 0
    if (used < LB_WINDOW) __builtin_amdgcn_s_sleep(1);
  }
  __hip_atomic_store(own, lb_word(LB_INCL, epoch, prefix + count), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  return prefix;
}

// The same with a whole wave: all 64 lanes call it with the same arguments
// and each lane loads one earlier word per step, so a step covers 64 tiles.
// A single lane covering LB_WINDOW words per memory round trip caps how many
// tiles a pass can finish per microsecond; passes over thousands of short
// tiles run into that cap.  Returns the prefix in every lane.
__device__ inline uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

__device__ inline uint32_t lookback_exclusive_wave(uint64_t* words, uint64_t stride,
                                                   uint64_t tile, uint32_t count,
                                                   uint32_t epoch) {
  epoch &= 0x3FFFFFFFu;
  const uint32_t lane = __lane_id();
  uint64_t* own = words + tile * stride;
  if (tile == 0) {
    if (lane == 0)
      __hip_atomic_store(own, lb_word(LB_INCL, epoch, count), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    return 0;
  }
  if (lane == 0)
    __hip_atomic_store(own, lb_word(LB_AGG, epoch, count), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  uint32_t prefix = 0;
  int64_t t = (int64_t)tile - 1;  // the same in every lane
  while (true) {
    const int64_t ti = t - (int64_t)lane;
    // lanes past tile 0 read nothing: tile 0's inclusive word ends the walk
    // before them
    const uint64_t x = ti >= 0 ? __hip_atomic_load(words + (uint64_t)ti * stride,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                               : 0ull;
    const bool valid = ti >= 0 && (x & LB_FLAGS) != 0 &&
                       (uint32_t)((x >> 32) & 0x3FFFFFFFu) == epoch;
    const uint64_t bad = __ballot(!valid), incl = __ballot(valid && (x & LB_FLAGS) == LB_INCL);
    const uint32_t first_bad = bad ? (uint32_t)__builtin_ctzll(bad) : 64u;
    const uint32_t first_incl = incl ? (uint32_t)__builtin_ctzll(incl) : 64u;
    if (first_incl < first_bad) {
      // every word up to the nearest inclusive one is published
      prefix += wave_sum_u32(lane <= first_incl ? (uint32_t)x : 0u);
      break;
    }
    if (first_bad > 0) {
      // the published aggregates before the first unpublished word
      prefix += wave_sum_u32(lane < first_bad ? (uint32_t)x : 0u);
      t -= first_bad;  // tile 0 publishes an inclusive word, so t stays >= 0
    } else {
      __builtin_amdgcn_s_sleep(1);
    }
  }
  if (lane == 0)
    __hip_atomic_store(own, lb_word(LB_INCL, epoch, prefix + count), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  return prefix;
}

}  // namespace svo
