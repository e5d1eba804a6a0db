// k_sort.hip -- LSD radix sort of voxel keys: 8-bit digits, one read for all
// digit histograms, then one kernel per digit pass (onesweep).
//
// Replaces the reference's dense flag array + exclusive scan + compaction
// (SVO.cu:367-405), which is O(res^3), with a sort of the emitted hit keys,
// O(hits).  The sorted order equals the reference's index order of the dense
// array (keys are the Morton codes).
//
// The key count n is read from device memory, so no launch shape depends on
// the hit count.  Kernels:
//   k_sort_clear  zero the histograms, tile tickets and the status words of
//                 the tiles the n keys occupy
//   k_upsweep     digit histograms of every pass in one read of the keys
//   k_scan_hist   exclusive scan of each pass's histogram -> digit offsets
//   k_onesweep    per pass: persistent blocks claim 4096-key tiles by atomic
//                 ticket, rank the keys stably with wave64 match masks
//                 (8 ballots per 64 keys), publish per-digit tile counts and
//                 find their exclusive prefix over earlier tiles by
//                 decoupled look-back, then write through an LDS reorder so
//                 the stores are contiguous within each digit run.
// Stable, so the passes compose into a full LSD sort.  HBM traffic for n keys
// of B bytes and P passes: n*B (upsweep) + P * 2*n*B.
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

__global__ __launch_bounds__(BLOCK) void k_sort_clear(uint64_t* status, uint32_t* hist,
                                                      uint32_t* tickets,
                                                      const unsigned long long* n_ptr) {
  const uint64_t tiles = (*n_ptr + TILE - 1) / TILE;
  const uint64_t words = tiles * RADIX;
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < words; i += stride)
    status[i] = 0;
  if (blockIdx.x == 0) {
    for (int i = threadIdx.x; i < MAX_PASSES * RADIX; i += BLOCK) hist[i] = 0;
    if (threadIdx.x < MAX_PASSES) tickets[threadIdx.x] = 0;
  }
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

// One digit pass.  Blocks loop taking tickets; a ticket holder is always a
// running block, so the look-back of tile t only waits on tiles < t whose
// blocks are already past their ticket and publish without waiting on any
// later tile: every wait ends.  Tickets at or past the last tile exit.
template <typename K>
__global__ __launch_bounds__(BLOCK) void k_onesweep(const K* in, K* out,
                                                    const unsigned long long* n_ptr, int shift,
                                                    uint32_t dm, const uint32_t* offs,
                                                    uint64_t* status,
                                                    uint32_t* ticket, uint32_t epoch) {
  __shared__ uint32_t wcnt[WAVES][RADIX];
  __shared__ uint32_t dbase[RADIX];
  __shared__ uint32_t tstart[RADIX];
  __shared__ uint32_t tmp[4];
  __shared__ uint32_t s_tile;
  __shared__ K sorted[TILE];
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t lane = __lane_id();
  const uint64_t n = *n_ptr;
  const uint32_t d_off = offs[threadIdx.x];
  while (true) {
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    for (int i = lane; i < RADIX; i += 64) wcnt[w][i] = 0;
    __syncthreads();
    const uint64_t tile = s_tile;
    const uint64_t tb = tile * TILE;
    if (tb >= n) return;  // uniform over the block
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
      const uint32_t prefix = lookback_exclusive(status + d, RADIX, tile, cnt, epoch);
      dbase[d] = d_off + prefix;
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
  return (size_t)tiles * RADIX * sizeof(uint64_t) + 2 * MAX_PASSES * RADIX * sizeof(uint32_t) +
         256;
}

template <typename K>
void launch_radix_sort(K* a, K* b, const unsigned long long* n, uint64_t cap, int bits,
                       uint32_t blocks, void* tmp, K** result, hipStream_t s) {
  const uint64_t tiles = (std::max<uint64_t>(cap, 1) + TILE - 1) / TILE;
  uint64_t* status = (uint64_t*)tmp;
  uint32_t* hist = (uint32_t*)(status + tiles * RADIX);
  uint32_t* offs = hist + MAX_PASSES * RADIX;
  uint32_t* tickets = offs + MAX_PASSES * RADIX;
  const int passes = std::min((bits + 7) / 8, MAX_PASSES);
  K* src = a;
  K* dst = b;
  if (passes > 0) {
    const uint32_t cb = (uint32_t)std::min<uint64_t>(2048, (tiles * RADIX + BLOCK - 1) / BLOCK);
    hipLaunchKernelGGL(k_sort_clear, dim3(std::max<uint32_t>(cb, 1)), dim3(BLOCK), 0, s, status,
                       hist, tickets, n);
    const uint64_t chunks = (cap + UP_KEYS - 1) / UP_KEYS;
    const uint32_t ub =
        (uint32_t)std::min<uint64_t>(2048, std::max<uint64_t>(1, (chunks + BLOCK - 1) / BLOCK));
    hipLaunchKernelGGL(k_upsweep<K>, dim3(ub), dim3(BLOCK), 0, s, (const K*)a, n, bits, hist);
    hipLaunchKernelGGL(k_scan_hist, dim3(1), dim3(BLOCK), 0, s, (const uint32_t*)hist, passes,
                       offs);
  }
  const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(tiles, blocks));
  for (int p = 0; p < passes; ++p) {
    hipLaunchKernelGGL(k_onesweep<K>, dim3(grid), dim3(BLOCK), 0, s, (const K*)src, dst, n,
                       8 * p, digit_mask(bits, 8 * p), (const uint32_t*)(offs + p * RADIX),
                       status, tickets + p,
                       (uint32_t)(p + 1));
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
