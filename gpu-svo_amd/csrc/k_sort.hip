// k_sort.hip -- LSD radix sort of voxel keys, 8-bit digits, wave64 ranking.
//
// Replaces the reference's dense flag array + exclusive scan + compaction
// (SVO.cu:367-405), which is O(res^3), with a sort of the emitted hit keys,
// O(hits).  The sorted order equals the reference's index order of the dense
// array (keys are the Morton codes).
//
// Fixed grid of G blocks; block b owns the contiguous key range
// [b*ceil(n/G), (b+1)*ceil(n/G)) where n is read from device memory, so the
// launch shapes do not depend on the hit count.  Per digit pass:
//   k_count    per-block digit histogram (per-wave LDS histograms)
//   k_rowscan  one block per digit: exclusive scan over the G blocks
//   k_scatter  per 4096-key tile: stable ranking with wave64 match masks
//              (8 ballots per 64 keys), LDS reorder by digit, then writes
//              that are contiguous within each digit run.
// Stable, so the passes compose into a full LSD sort.
#include "svo_internal.h"
#include "svo_scan.h"

namespace svo {
namespace {

constexpr int BLOCK = 256;
constexpr int WAVES = BLOCK / 64;
constexpr int RADIX = 256;
constexpr int KPT = 16;                       // keys per lane per tile
constexpr int TILE = WAVES * 64 * KPT;        // 4096

__device__ inline void block_range(uint64_t n, uint32_t G, uint32_t b, uint64_t& beg,
                                   uint64_t& end) {
  const uint64_t chunk = (n + G - 1) / G;
  beg = (uint64_t)b * chunk;
  if (beg > n) beg = n;
  end = beg + chunk;
  if (end > n) end = n;
}

template <typename K>
__global__ __launch_bounds__(BLOCK) void k_count(const K* in, const unsigned long long* n_ptr,
                                                 int shift, uint32_t G, uint32_t* counts) {
  __shared__ uint32_t hist[WAVES][RADIX];
  const uint32_t w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < WAVES * RADIX; i += BLOCK) (&hist[0][0])[i] = 0;
  __syncthreads();
  uint64_t beg, end;
  block_range(*n_ptr, G, blockIdx.x, beg, end);
  for (uint64_t i = beg + threadIdx.x; i < end; i += BLOCK) {
    const uint32_t d = (uint32_t)(in[i] >> shift) & (RADIX - 1);
    atomicAdd(&hist[w][d], 1u);
  }
  __syncthreads();
  const uint32_t d = threadIdx.x;
  uint32_t s = 0;
  for (int k = 0; k < WAVES; ++k) s += hist[k][d];
  counts[(uint64_t)d * G + blockIdx.x] = s;
}

// Block d: exclusive scan of counts[d][0..G) in place, row total to row_tot[d].
__global__ __launch_bounds__(BLOCK) void k_rowscan(uint32_t* counts, uint32_t G,
                                                   uint32_t* row_tot) {
  __shared__ uint32_t tmp[4];
  uint32_t* row = counts + (uint64_t)blockIdx.x * G;
  const uint32_t per = (G + BLOCK - 1) / BLOCK;
  const uint32_t b0 = threadIdx.x * per;
  uint32_t local = 0;
  for (uint32_t k = 0; k < per; ++k)
    if (b0 + k < G) local += row[b0 + k];
  uint32_t total;
  uint32_t run = block_excl_scan256(local, tmp, total);
  for (uint32_t k = 0; k < per; ++k)
    if (b0 + k < G) {
      const uint32_t c = row[b0 + k];
      row[b0 + k] = run;
      run += c;
    }
  if (threadIdx.x == 0) row_tot[blockIdx.x] = total;
}

template <typename K>
__global__ __launch_bounds__(BLOCK) void k_scatter(const K* in, K* out,
                                                   const unsigned long long* n_ptr, int shift,
                                                   uint32_t G, const uint32_t* counts,
                                                   const uint32_t* row_tot) {
  __shared__ uint32_t wcnt[WAVES][RADIX];
  __shared__ uint32_t dbase[RADIX];
  __shared__ uint32_t tstart[RADIX];
  __shared__ uint32_t tcount[RADIX];
  __shared__ uint32_t tmp[4];
  __shared__ K sorted[TILE];
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t lane = __lane_id();
  uint64_t beg, end;
  block_range(*n_ptr, G, blockIdx.x, beg, end);
  {
    uint32_t total;
    const uint32_t ex = block_excl_scan256(row_tot[threadIdx.x], tmp, total);
    dbase[threadIdx.x] = ex + counts[(uint64_t)threadIdx.x * G + blockIdx.x];
  }
  for (uint64_t tb = beg; tb < end; tb += TILE) {
    const uint32_t tn = (uint32_t)((end - tb) < TILE ? (end - tb) : TILE);
    for (int i = lane; i < RADIX; i += 64) wcnt[w][i] = 0;
    __syncthreads();
    K keys[KPT];
    uint32_t rank[KPT];
#pragma unroll
    for (int it = 0; it < KPT; ++it) {
      const uint32_t pos = w * (KPT * 64) + it * 64 + lane;
      const bool valid = pos < tn;
      const K key = valid ? in[tb + pos] : (K)0;
      const uint32_t d = (uint32_t)(key >> shift) & (RADIX - 1);
      uint64_t eq = __ballot(valid);
#pragma unroll
      for (int bit = 0; bit < 8; ++bit) {
        const bool b = (d >> bit) & 1u;
        const uint64_t m = __ballot(b);
        eq &= b ? m : ~m;
      }
      const uint32_t leader = valid ? (uint32_t)__builtin_ctzll(eq) : 0u;
      uint32_t old = 0;
      if (valid && lane == leader) {
        old = wcnt[w][d];
        wcnt[w][d] = old + (uint32_t)__popcll(eq);
      }
      old = __shfl(old, (int)leader);
      rank[it] = old + lanes_below(eq);
      keys[it] = key;
    }
    __syncthreads();
    {
      const uint32_t d = threadIdx.x;
      uint32_t s = 0;
#pragma unroll
      for (int k = 0; k < WAVES; ++k) {
        const uint32_t c = wcnt[k][d];
        wcnt[k][d] = s;
        s += c;
      }
      tcount[d] = s;
      uint32_t total;
      tstart[d] = block_excl_scan256(s, tmp, total);
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < KPT; ++it) {
      const uint32_t pos = w * (KPT * 64) + it * 64 + lane;
      if (pos < tn) {
        const uint32_t d = (uint32_t)(keys[it] >> shift) & (RADIX - 1);
        sorted[tstart[d] + wcnt[w][d] + rank[it]] = keys[it];
      }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < tn; i += BLOCK) {
      const K key = sorted[i];
      const uint32_t d = (uint32_t)(key >> shift) & (RADIX - 1);
      out[(uint64_t)dbase[d] + (i - tstart[d])] = key;
    }
    __syncthreads();
    dbase[threadIdx.x] += tcount[threadIdx.x];
    __syncthreads();
  }
}

}  // namespace

size_t radix_sort_tmp_bytes(uint32_t blocks) {
  return (size_t)RADIX * blocks * sizeof(uint32_t) + RADIX * sizeof(uint32_t);
}

template <typename K>
void launch_radix_sort(K* a, K* b, const unsigned long long* n, uint64_t cap, int bits,
                       uint32_t blocks, void* tmp, K** result, hipStream_t s) {
  (void)cap;
  uint32_t* counts = (uint32_t*)tmp;
  uint32_t* row_tot = counts + (size_t)RADIX * blocks;
  K* src = a;
  K* dst = b;
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_count<K>, dim3(blocks), dim3(BLOCK), 0, s, (const K*)src, n, shift,
                       blocks, counts);
    hipLaunchKernelGGL(k_rowscan, dim3(RADIX), dim3(BLOCK), 0, s, counts, blocks, row_tot);
    hipLaunchKernelGGL(k_scatter<K>, dim3(blocks), dim3(BLOCK), 0, s, (const K*)src, dst, n,
                       shift, blocks, (const uint32_t*)counts, (const uint32_t*)row_tot);
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
