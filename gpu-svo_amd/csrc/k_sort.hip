// k_sort.hip -- LSD radix sort of voxel keys: 8-bit digits, one read for all
// digit histograms, then three kernels per digit pass (reduce, scan, scatter).
//
// Replaces the reference's dense flag array + exclusive scan + compaction
// (SVO.cu:367-405), which is O(res^3), with a sort of the emitted hit keys,
// O(hits).  The sorted order equals the reference's index order of the dense
// array (keys are the Morton codes).
//
// The key count n is read from device memory, so no launch shape depends on
// the hit count.  Kernels:
//   k_sort_clear  zero the histograms
//   k_upsweep     digit histograms of every pass in one read of the keys
//   k_scan_hist   exclusive scan of each pass's histogram -> digit offsets
//   per pass:
//   k_tile_hist   digit counts of each 4096-key tile, stored digit-major
//   k_tile_scan   per digit, exclusive scan over the tiles plus the digit's
//                 offset -> where each tile's run of each digit starts
//   k_scatter     rank the tile's keys stably with wave64 match masks (8
//                 ballots per 64 keys) and write them through an LDS reorder
//                 so the stores are contiguous within each digit run
// Stable, so the passes compose into a full LSD sort.  HBM traffic for n keys
// of B bytes and P passes: n*B (upsweep) + P * 3*n*B, plus 2 KB of counts per
// tile and pass.  A single-pass scatter with decoupled look-back reads and
// writes the keys once per pass, but over thousands of tiles its look-back
// chain (and the 256 digits' status words each tile walks) bounds the pass;
// here no tile waits on another.
#include <algorithm>

#include "svo_internal.h"
#include "svo_scan.h"

namespace svo {
namespace {

constexpr int BLOCK = 256;
constexpr int WAVES = BLOCK / 64;
constexpr int RADIX = 256;
constexpr int MAX_PASSES = 8;
constexpr int KPT = 16;                       // keys per lane per tile
constexpr int TILE = WAVES * 64 * KPT;        // 4096
constexpr int UP_KEYS = 16;                   // consecutive keys per upsweep thread

// Digit of the pass at `shift` covers bits [shift, min(shift + 8, bits)).
__host__ __device__ inline uint32_t digit_mask(int bits, int shift) {
  const int w = bits - shift < 8 ? bits - shift : 8;
  return (1u << w) - 1u;
}

__global__ __launch_bounds__(BLOCK) void k_sort_clear(uint32_t* hist) {
  for (int i = threadIdx.x; i < MAX_PASSES * RADIX; i += BLOCK) hist[i] = 0;
}

// Each thread owns UP_KEYS consecutive keys and adds runs of equal digits to
// the block histogram: high digits (mesh index, coarse cells) come in long
// runs because nearby triangles emit nearby keys, which keeps LDS atomics
// off the hot bins.
template <typename K>
__global__ __launch_bounds__(BLOCK) void k_upsweep(const K* in, const unsigned long long* n_ptr,
                                                   int bits, uint32_t* hist) {
  const int passes = (bits + 7) / 8;
  __shared__ uint32_t h[MAX_PASSES][RADIX];
  for (int i = threadIdx.x; i < MAX_PASSES * RADIX; i += BLOCK) (&h[0][0])[i] = 0;
  __syncthreads();
  const uint64_t n = *n_ptr;
  const uint64_t chunks = (n + UP_KEYS - 1) / UP_KEYS;
  const bool vec = ((uintptr_t)in & 15) == 0;
  for (uint64_t c = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; c < chunks;
       c += (uint64_t)gridDim.x * BLOCK) {
    const uint64_t b = c * UP_KEYS;
    const uint32_t cnt = (uint32_t)((n - b) < UP_KEYS ? (n - b) : UP_KEYS);
    K kk[UP_KEYS];
    load_consecutive(in + b, vec, cnt, kk);
    for (int p = 0; p < passes; ++p) {
      const int sh = 8 * p;
      const uint32_t dm = digit_mask(bits, sh);
      uint32_t last = (uint32_t)(kk[0] >> sh) & dm, run = 1;
#pragma unroll
      for (int j = 1; j < UP_KEYS; ++j) {
        if ((uint32_t)j < cnt) {
          const uint32_t d = (uint32_t)(kk[j] >> sh) & dm;
          if (d == last) {
            ++run;
          } else {
            atomicAdd(&h[p][last], run);
            last = d;
            run = 1;
          }
        }
      }
      atomicAdd(&h[p][last], run);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < passes * RADIX; i += BLOCK) {
    const uint32_t v = (&h[0][0])[i];
    if (v) atomicAdd(hist + i, v);
  }
}

__global__ __launch_bounds__(BLOCK) void k_scan_hist(const uint32_t* hist, int passes,
                                                     uint32_t* offs) {
  __shared__ uint32_t tmp[4];
  for (int p = 0; p < passes; ++p) {
    uint32_t total;
    offs[p * RADIX + threadIdx.x] =
        block_excl_scan256(hist[p * RADIX + threadIdx.x], tmp, total);
  }
}

// Digit counts of tiles blockIdx.x, + gridDim.x, ...: counts[d * tiles + t].
template <typename K>
__global__ __launch_bounds__(BLOCK) void k_tile_hist(const K* in, const unsigned long long* n_ptr,
                                                     int shift, uint32_t dm, uint32_t* counts,
                                                     uint64_t tiles) {
  __shared__ uint32_t h[WAVES][RADIX];
  const uint32_t w = threadIdx.x >> 6;
  const uint64_t n = *n_ptr;
  const uint64_t n_tiles = (n + TILE - 1) / TILE;
  const bool vec = ((uintptr_t)in & 15) == 0;
  for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    for (int i = __lane_id(); i < RADIX; i += 64) h[w][i] = 0;
    __syncthreads();
    // thread: KPT consecutive keys of the tile; runs of equal digits are
    // added at once
    const uint64_t b = tile * TILE + (uint64_t)threadIdx.x * KPT;
    const uint32_t cnt = b < n ? (uint32_t)((n - b) < KPT ? (n - b) : KPT) : 0u;
    K kk[KPT];
    load_consecutive(in + (b < n ? b : 0), vec && b < n, cnt, kk);
    if (cnt) {
      uint32_t last = (uint32_t)(kk[0] >> shift) & dm, run = 1;
#pragma unroll
      for (int j = 1; j < KPT; ++j) {
        if ((uint32_t)j < cnt) {
          const uint32_t d = (uint32_t)(kk[j] >> shift) & dm;
          if (d == last) {
            ++run;
          } else {
            atomicAdd(&h[w][last], run);
            last = d;
            run = 1;
          }
        }
      }
      atomicAdd(&h[w][last], run);
    }
    __syncthreads();
    {
      const uint32_t d = threadIdx.x;
      uint32_t c = 0;
#pragma unroll
      for (int k = 0; k < WAVES; ++k) c += h[k][d];
      counts[(uint64_t)d * tiles + tile] = c;
    }
    __syncthreads();
  }
}

// Block d: exclusive scan of counts[d * tiles + 0 .. n_tiles) in place, plus
// the digit's offset in the pass's output.
__global__ __launch_bounds__(BLOCK) void k_tile_scan(uint32_t* counts,
                                                     const unsigned long long* n_ptr,
                                                     uint64_t tiles, const uint32_t* offs) {
  __shared__ uint32_t tmp[4];
  const uint64_t n_tiles = (*n_ptr + TILE - 1) / TILE;
  uint32_t* c = counts + (uint64_t)blockIdx.x * tiles;
  uint32_t run = offs[blockIdx.x];
  constexpr int PER = 16;
  for (uint64_t t0 = 0; t0 < n_tiles; t0 += (uint64_t)BLOCK * PER) {
    const uint64_t i0 = t0 + (uint64_t)threadIdx.x * PER;
    uint32_t v[PER], sum = 0;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      v[j] = i0 + j < n_tiles ? c[i0 + j] : 0u;
      sum += v[j];
    }
    uint32_t total;
    uint32_t p = run + block_excl_scan256(sum, tmp, total);
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      if (i0 + j < n_tiles) c[i0 + j] = p;
      p += v[j];
    }
    run += total;
    __syncthreads();  // tmp is reused by the next chunk's scan
  }
}

// One digit pass over tiles blockIdx.x, + gridDim.x, ...: the tile's keys
// are ranked stably, reordered in LDS by digit and written to where
// k_tile_scan put the tile's run of each digit.
template <typename K>
__global__ __launch_bounds__(BLOCK) void k_scatter(const K* in, K* out,
                                                   const unsigned long long* n_ptr, int shift,
                                                   uint32_t dm, const uint32_t* starts,
                                                   uint64_t tiles) {
  __shared__ uint32_t wcnt[WAVES][RADIX];
  __shared__ uint32_t dbase[RADIX];
  __shared__ uint32_t tstart[RADIX];
  __shared__ uint32_t tmp[4];
  __shared__ K sorted[TILE];
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t lane = __lane_id();
  const uint64_t n = *n_ptr;
  const uint64_t n_tiles = (n + TILE - 1) / TILE;
  for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    for (int i = lane; i < RADIX; i += 64) wcnt[w][i] = 0;
    dbase[threadIdx.x] = starts[(uint64_t)threadIdx.x * tiles + tile];
    __syncthreads();
    const uint64_t tb = tile * TILE;
    const uint32_t tn = (uint32_t)((n - tb) < TILE ? (n - tb) : TILE);
    // loads first, all in flight together (indices clamped, so unconditional)
    K keys[KPT];
#pragma unroll
    for (int it = 0; it < KPT; ++it) {
      const uint32_t pos = w * (KPT * 64) + it * 64 + lane;
      const K v = in[tb + (pos < tn ? pos : tn - 1)];
      keys[it] = pos < tn ? v : (K)0;
    }
    // stable rank within the wave: match masks per 64 keys; the group leader
    // reserves the group's slots in the wave's digit counter (LDS atomics
    // return the old counts; the 16 reservations are waited on together)
    uint32_t rank[KPT], lead[KPT], base[KPT];
#pragma unroll
    for (int it = 0; it < KPT; ++it) {
      const uint32_t pos = w * (KPT * 64) + it * 64 + lane;
      const bool valid = pos < tn;
      const uint32_t d = (uint32_t)(keys[it] >> shift) & dm;
      uint64_t eq = __ballot(valid);
#pragma unroll
      for (int bit = 0; bit < 8; ++bit) {
        // lanes whose bit equals this lane's: ~(ballot ^ (bit ? ~0 : 0))
        const uint32_t sx = (uint32_t)((int32_t)(d << (31 - bit)) >> 31);
        const uint64_t m = __ballot(sx != 0);
        eq &= ~(m ^ (((uint64_t)sx << 32) | sx));
      }
      lead[it] = valid ? (uint32_t)__builtin_ctzll(eq) : 0u;
      rank[it] = lanes_below(eq);
      base[it] = 0;
      if (valid && lane == lead[it]) base[it] = atomicAdd(&wcnt[w][d], (uint32_t)__popcll(eq));
    }
#pragma unroll
    for (int it = 0; it < KPT; ++it) rank[it] += __shfl(base[it], (int)lead[it]);
    __syncthreads();
    {
      const uint32_t d = threadIdx.x;
      uint32_t cnt = 0;
#pragma unroll
      for (int k = 0; k < WAVES; ++k) {
        const uint32_t c = wcnt[k][d];
        wcnt[k][d] = cnt;
        cnt += c;
      }
      uint32_t total;
      tstart[d] = block_excl_scan256(cnt, tmp, total);
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < KPT; ++it) {
      const uint32_t pos = w * (KPT * 64) + it * 64 + lane;
      if (pos < tn) {
        const uint32_t d = (uint32_t)(keys[it] >> shift) & dm;
        sorted[tstart[d] + wcnt[w][d] + rank[it]] = keys[it];
      }
    }
    __syncthreads();
    // write out in tile order: consecutive threads store consecutive keys of
    // a digit run
    K sk[KPT];
#pragma unroll
    for (int it = 0; it < KPT; ++it) {
      const uint32_t i = it * BLOCK + threadIdx.x;
      sk[it] = sorted[i < tn ? i : 0];
    }
#pragma unroll
    for (int it = 0; it < KPT; ++it) {
      const uint32_t i = it * BLOCK + threadIdx.x;
      const uint32_t d = (uint32_t)(sk[it] >> shift) & dm;
      const uint64_t o = (uint64_t)dbase[d] + (i - tstart[d]);
      if (i < tn) out[o] = sk[it];
    }
    __syncthreads();
  }
}

}  // namespace

size_t radix_sort_tmp_bytes(uint64_t cap) {
  const uint64_t tiles = (std::max<uint64_t>(cap, 1) + TILE - 1) / TILE;
  return (size_t)tiles * RADIX * sizeof(uint32_t) + 2 * MAX_PASSES * RADIX * sizeof(uint32_t) +
         256;
}

template <typename K>
void launch_radix_sort(K* a, K* b, const unsigned long long* n, uint64_t cap, int bits,
                       uint32_t blocks, void* tmp, K** result, hipStream_t s) {
  const uint64_t tiles = (std::max<uint64_t>(cap, 1) + TILE - 1) / TILE;
  uint32_t* counts = (uint32_t*)tmp;
  uint32_t* hist = counts + tiles * RADIX;
  uint32_t* offs = hist + MAX_PASSES * RADIX;
  const int passes = std::min((bits + 7) / 8, MAX_PASSES);
  K* src = a;
  K* dst = b;
  if (passes > 0) {
    hipLaunchKernelGGL(k_sort_clear, dim3(1), dim3(BLOCK), 0, s, hist);
    const uint64_t chunks = (cap + UP_KEYS - 1) / UP_KEYS;
    const uint32_t ub =
        (uint32_t)std::min<uint64_t>(2048, std::max<uint64_t>(1, (chunks + BLOCK - 1) / BLOCK));
    hipLaunchKernelGGL(k_upsweep<K>, dim3(ub), dim3(BLOCK), 0, s, (const K*)a, n, bits, hist);
    hipLaunchKernelGGL(k_scan_hist, dim3(1), dim3(BLOCK), 0, s, (const uint32_t*)hist, passes,
                       offs);
  }
  // tiles are independent: at most `blocks` workgroups, each taking tiles
  // blockIdx.x, + gridDim.x, ...
  const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(tiles, blocks));
  for (int p = 0; p < passes; ++p) {
    const uint32_t dm = digit_mask(bits, 8 * p);
    hipLaunchKernelGGL(k_tile_hist<K>, dim3(grid), dim3(BLOCK), 0, s, (const K*)src, n, 8 * p,
                       dm, counts, tiles);
    hipLaunchKernelGGL(k_tile_scan, dim3(RADIX), dim3(BLOCK), 0, s, counts, n, tiles,
                       (const uint32_t*)(offs + p * RADIX));
    hipLaunchKernelGGL(k_scatter<K>, dim3(grid), dim3(BLOCK), 0, s, (const K*)src, dst, n,
                       8 * p, dm, (const uint32_t*)counts, tiles);
    K* t = src;
    src = dst;
    dst = t;
  }
  *result = src;
}

template void launch_radix_sort<uint32_t>(uint32_t*, uint32_t*, const unsigned long long*,
                                          uint64_t, int, uint32_t, void*, uint32_t**,
                                          hipStream_t);
template void launch_radix_sort<uint64_t>(uint64_t*, uint64_t*, const unsigned long long*,
                                          uint64_t, int, uint32_t, void*, uint64_t**,
                                          hipStream_t);

}  // namespace svo
