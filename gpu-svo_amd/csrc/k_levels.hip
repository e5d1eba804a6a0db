// k_levels.hip -- sorted unique voxels and the bottom-up octree levels.
//
// Reference: the level loop of SparseVoxelOctree::create (SVO.cu:383-495).
// Per level it flags parent-group heads (cpNumNodes, SVO.cu:221-235), scans
// them (SVO.cu:433-437), places node i at sum[i] + (code & 7) (createNode_*,
// SVO.cu:237-309), fills the empty siblings (createRemainNode, 311-339) and
// stops at the first level holding a single code (487-490).
//
// Here every level is one scan over the level's sorted unique keys:
//   group head  = first key of its parent group (parent differs from the
//                 previous key's parent; composite keys keep meshes apart),
//   root        = the only key of its mesh at this level (the reference's
//                 single-code level), which ends that mesh's tree,
//   output      = next level's keys (the group parents, already sorted and
//                 unique) and, per group, the index of its first child.
// Level 1 is fused with the de-duplication of the sorted hit keys.  Node
// records are written later from these arrays (k_attrs.hip), once the node
// counts are known.  Each scan is count -> block-offset scan -> write over a
// fixed grid, with the element count read from device memory.
#include "svo_internal.h"
#include "svo_scan.h"

namespace svo {
namespace {

constexpr int BLOCK = 256;
constexpr int EPT = 8;  // elements per thread per tile
constexpr int TILE = BLOCK * EPT;

__device__ inline void block_range(uint64_t n, uint32_t G, uint32_t b, uint64_t& beg,
                                   uint64_t& end) {
  const uint64_t chunk = (n + G - 1) / G;
  beg = (uint64_t)b * chunk;
  if (beg > n) beg = n;
  end = beg + chunk;
  if (end > n) end = n;
}

// Loads keys [tb - 1, tb + tn] into s[0 .. tn + 1] with coalesced reads, so
// that every element of the tile sees its neighbours in LDS.
template <typename K>
__device__ inline void load_tile(const K* src, uint64_t n, uint64_t tb, uint32_t tn, K* s) {
  __syncthreads();  // the previous tile is no longer read
  for (uint32_t j = threadIdx.x; j < tn + 2; j += BLOCK) {
    const uint64_t gi = tb + j;  // element gi - 1
    s[j] = (gi >= 1 && gi - 1 < n) ? src[gi - 1] : (K)0;
  }
  __syncthreads();
}

template <typename K>
__device__ inline uint32_t first_flags_v(K prev, K cur, bool has_prev, KeyLayout<K> kl) {
  if (!has_prev) return 3u;
  if (cur == prev) return 0u;
  return 1u | (kl.parent(cur) != kl.parent(prev) ? 2u : 0u);
}

template <typename K>
__device__ inline uint32_t level_flags_v(K prev, K cur, K next, bool has_prev, bool has_next,
                                         KeyLayout<K> kl, const MeshTreeState* state) {
  const uint32_t m = kl.mesh(cur);
  if (state[m].depth != 0) return 0u;
  const bool first = !has_prev || kl.mesh(prev) != m;
  const bool last = !has_next || kl.mesh(next) != m;
  if (first && last) return 4u;
  if (first) return 2u;
  return kl.parent(cur) != kl.parent(prev) ? 2u : 0u;
}

// ---------------------------------------------------------------- level 1
// flags of sorted hit key i: bit0 = first copy of its key, bit1 = parent-group head
template <typename K>
__device__ inline uint32_t first_flags(const K* k, uint64_t i, KeyLayout<K> kl) {
  const K cur = k[i];
  if (i == 0) return 3u;
  const K prev = k[i - 1];
  if (cur == prev) return 0u;
  return 1u | (kl.parent(cur) != kl.parent(prev) ? 2u : 0u);
}

template <typename K>
__global__ __launch_bounds__(BLOCK) void k_first_count(const K* sorted,
                                                       const unsigned long long* n_ptr,
                                                       KeyLayout<K> kl, uint32_t G,
                                                       uint64_t* partial) {
  __shared__ uint64_t tmp[4];
  uint64_t beg, end;
  block_range(*n_ptr, G, blockIdx.x, beg, end);
  uint64_t nu = 0, ng = 0;
  for (uint64_t i = beg + threadIdx.x; i < end; i += BLOCK) {
    const uint32_t f = first_flags(sorted, i, kl);
    nu += f & 1u;
    ng += f >> 1;
  }
  uint64_t tu, tg;
  block_excl_scan256(nu, tmp, tu);
  block_excl_scan256(ng, tmp, tg);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = tu;
    partial[2 * blockIdx.x + 1] = tg;
  }
}

// Exclusive scan of G pairs (single block); totals to *out0 / *out1 when set.
__global__ __launch_bounds__(BLOCK) void k_pair_scan(uint64_t* partial, uint32_t G,
                                                     unsigned long long* out0,
                                                     unsigned long long* out1) {
  __shared__ uint64_t tmp[4];
  const uint32_t per = (G + BLOCK - 1) / BLOCK;
  const uint32_t b0 = threadIdx.x * per;
  uint64_t s0 = 0, s1 = 0;
  for (uint32_t k = 0; k < per; ++k)
    if (b0 + k < G) {
      s0 += partial[2 * (b0 + k)];
      s1 += partial[2 * (b0 + k) + 1];
    }
  uint64_t t0, t1;
  uint64_t r0 = block_excl_scan256(s0, tmp, t0);
  uint64_t r1 = block_excl_scan256(s1, tmp, t1);
  for (uint32_t k = 0; k < per; ++k)
    if (b0 + k < G) {
      const uint64_t c0 = partial[2 * (b0 + k)], c1 = partial[2 * (b0 + k) + 1];
      partial[2 * (b0 + k)] = r0;
      partial[2 * (b0 + k) + 1] = r1;
      r0 += c0;
      r1 += c1;
    }
  if (threadIdx.x == 0) {
    if (out0) *out0 = t0;
    if (out1) *out1 = t1;
  }
}

template <typename K>
__global__ __launch_bounds__(BLOCK) void k_first_write(const K* sorted,
                                                       const unsigned long long* n_ptr,
                                                       KeyLayout<K> kl, uint32_t G,
                                                       const uint64_t* partial, K* elem1,
                                                       uint32_t* group1, K* elem2,
                                                       uint32_t* first_child2, LevelSpan* spans,
                                                       uint32_t n_levels) {
  __shared__ uint64_t tmp[4];
  __shared__ K sk[TILE + 2];
  const uint64_t n = *n_ptr;
  uint64_t beg, end;
  block_range(n, G, blockIdx.x, beg, end);
  uint64_t ubase = partial[2 * blockIdx.x], gbase = partial[2 * blockIdx.x + 1];
  for (uint64_t tb = beg; tb < end; tb += TILE) {
    const uint32_t tn = (uint32_t)((end - tb) < TILE ? (end - tb) : TILE);
    load_tile(sorted, n, tb, tn, sk);
    // blocked: thread t owns elements tb + t*EPT .. +EPT
    const uint32_t l0 = threadIdx.x * EPT;
    const uint64_t i0 = tb + l0;
    uint32_t f[EPT];
    uint64_t nu = 0, ng = 0;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const uint32_t l = l0 + e;
      f[e] = (l < tn) ? first_flags_v(sk[l], sk[l + 1], tb + l > 0, kl) : 0u;
      nu += f[e] & 1u;
      ng += f[e] >> 1;
    }
    uint64_t tu, tg;
    uint64_t ru = block_excl_scan256(nu, tmp, tu) + ubase;  // uniques before i0
    uint64_t rg = block_excl_scan256(ng, tmp, tg) + gbase;  // groups before i0
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const uint64_t i = i0 + e;
      if (i >= end) break;
      const K key = sk[l0 + e + 1];
      const uint32_t m = kl.mesh(key);
      if (f[e] & 1u) {
        if (f[e] & 2u) {
          elem2[rg] = kl.parent(key);
          first_child2[rg] = (uint32_t)ru;
          ++rg;
        }
        elem1[ru] = key;
        group1[ru] = (uint32_t)(rg - 1);
        // first element of its mesh
        if (i == 0 || kl.mesh(sk[l0 + e]) != m) {
          spans[(uint64_t)m * n_levels].elem_first = (uint32_t)ru;
          spans[(uint64_t)m * n_levels].group_first = (uint32_t)(rg - 1);
        }
        ++ru;
      }
      // last element (copy) of its mesh
      if (i + 1 == n || kl.mesh(sk[l0 + e + 2]) != m) {
        spans[(uint64_t)m * n_levels].elem_end = (uint32_t)ru;
        spans[(uint64_t)m * n_levels].group_end = (uint32_t)rg;
      }
    }
    ubase += tu;
    gbase += tg;
  }
}

__global__ void k_first_finalize(const LevelSpan* spans, uint32_t n_meshes, uint32_t n_levels,
                                 MeshTreeState* state) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n_meshes) return;
  const LevelSpan sp = spans[(uint64_t)m * n_levels];
  if (sp.elem_end - sp.elem_first == 1) {  // a single voxel: depth-1 tree
    state[m].depth = 1;
    state[m].root_elem = sp.elem_first;
  }
}

// ---------------------------------------------------------------- level l >= 2
// flags of element i: bit1 = group head, bit2 = root of its mesh, 0 = skipped
template <typename K>
__device__ inline uint32_t level_flags(const K* e, uint64_t i, uint64_t n, KeyLayout<K> kl,
                                       const MeshTreeState* state) {
  const K cur = e[i];
  const uint32_t m = kl.mesh(cur);
  if (state[m].depth != 0) return 0u;  // finished at a lower level (single voxel)
  const bool first = (i == 0) || kl.mesh(e[i - 1]) != m;
  const bool last = (i + 1 == n) || kl.mesh(e[i + 1]) != m;
  if (first && last) return 4u;
  if (first) return 2u;
  return kl.parent(cur) != kl.parent(e[i - 1]) ? 2u : 0u;
}

template <typename K>
__global__ __launch_bounds__(BLOCK) void k_level_count(const K* elem,
                                                       const unsigned long long* n_ptr,
                                                       KeyLayout<K> kl, const MeshTreeState* state,
                                                       uint32_t G, uint64_t* partial) {
  __shared__ uint64_t tmp[4];
  const uint64_t n = *n_ptr;
  uint64_t beg, end;
  block_range(n, G, blockIdx.x, beg, end);
  uint64_t ng = 0;
  for (uint64_t i = beg + threadIdx.x; i < end; i += BLOCK)
    ng += (level_flags(elem, i, n, kl, state) >> 1) & 1u;
  uint64_t tg;
  block_excl_scan256(ng, tmp, tg);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = 0;
    partial[2 * blockIdx.x + 1] = tg;
  }
}

template <typename K>
__global__ __launch_bounds__(BLOCK) void k_level_write(
    int level, const K* elem, const unsigned long long* n_ptr, KeyLayout<K> kl,
    MeshTreeState* state, uint32_t G, const uint64_t* partial, uint32_t* group_l, K* elem_next,
    uint32_t* first_child_next, LevelSpan* spans, uint32_t n_levels) {
  __shared__ uint64_t tmp[4];
  __shared__ K sk[TILE + 2];
  const uint64_t n = *n_ptr;
  uint64_t beg, end;
  block_range(n, G, blockIdx.x, beg, end);
  uint64_t gbase = partial[2 * blockIdx.x + 1];
  const uint32_t li = (uint32_t)(level - 1);
  for (uint64_t tb = beg; tb < end; tb += TILE) {
    const uint32_t tn = (uint32_t)((end - tb) < TILE ? (end - tb) : TILE);
    load_tile(elem, n, tb, tn, sk);
    const uint32_t l0 = threadIdx.x * EPT;
    const uint64_t i0 = tb + l0;
    uint32_t f[EPT];
    uint64_t ng = 0;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const uint32_t l = l0 + e;
      f[e] = (l < tn) ? level_flags_v(sk[l], sk[l + 1], sk[l + 2], tb + l > 0, tb + l + 1 < n,
                                      kl, state)
                      : 0u;
      ng += (f[e] >> 1) & 1u;
    }
    uint64_t tg;
    uint64_t rg = block_excl_scan256(ng, tmp, tg) + gbase;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const uint64_t i = i0 + e;
      if (i >= end) break;
      const K key = sk[l0 + e + 1];
      const uint32_t m = kl.mesh(key);
      if (f[e] & 4u) {  // the only key of its mesh at this level: the root
        state[m].depth = (uint32_t)level;
        state[m].root_elem = (uint32_t)i;
        group_l[i] = NONE;
        LevelSpan& sp = spans[(uint64_t)m * n_levels + li];
        sp.elem_first = (uint32_t)i;
        sp.elem_end = (uint32_t)i + 1;
        sp.group_first = sp.group_end = 0;
        continue;
      }
      if (state[m].depth != 0) {  // depth-1 tree carried over from level 1
        group_l[i] = NONE;
        continue;
      }
      if (f[e] & 2u) {
        elem_next[rg] = kl.parent(key);
        first_child_next[rg] = (uint32_t)i;
        ++rg;
      }
      group_l[i] = (uint32_t)(rg - 1);
      const bool first = (i == 0) || kl.mesh(sk[l0 + e]) != m;
      const bool last = (i + 1 == n) || kl.mesh(sk[l0 + e + 2]) != m;
      LevelSpan& sp = spans[(uint64_t)m * n_levels + li];
      if (first) {
        sp.elem_first = (uint32_t)i;
        sp.group_first = (uint32_t)(rg - 1);
      }
      if (last) {
        sp.elem_end = (uint32_t)i + 1;
        sp.group_end = (uint32_t)rg;
      }
    }
    gbase += tg;
  }
}

}  // namespace

size_t level_tmp_bytes(uint32_t blocks) { return (size_t)2 * blocks * sizeof(uint64_t); }

template <typename K>
void launch_level_first(const K* sorted, const unsigned long long* n_sorted, KeyLayout<K> kl,
                        LevelArrays l1, LevelArrays l2, unsigned long long* counts,
                        LevelSpan* spans1, uint32_t n_levels, uint32_t blocks, void* tmp,
                        hipStream_t s) {
  uint64_t* partial = (uint64_t*)tmp;
  hipLaunchKernelGGL(k_first_count<K>, dim3(blocks), dim3(BLOCK), 0, s, sorted, n_sorted, kl,
                     blocks, partial);
  hipLaunchKernelGGL(k_pair_scan, dim3(1), dim3(BLOCK), 0, s, partial, blocks, counts,
                     counts + 1);
  hipLaunchKernelGGL(k_first_write<K>, dim3(blocks), dim3(BLOCK), 0, s, sorted, n_sorted, kl,
                     blocks, (const uint64_t*)partial, (K*)l1.elem, l1.elem_group, (K*)l2.elem,
                     l2.first_child, spans1, n_levels);
}

void launch_level_first_finalize(const LevelSpan* spans, uint32_t n_meshes, uint32_t n_levels,
                                 MeshTreeState* state, hipStream_t s) {
  hipLaunchKernelGGL(k_first_finalize, dim3((n_meshes + 63) / 64), dim3(64), 0, s, spans,
                     n_meshes, n_levels, state);
}

template <typename K>
void launch_level(int level, KeyLayout<K> kl, LevelArrays cur, LevelArrays next,
                  unsigned long long* counts, LevelSpan* spans, uint32_t n_meshes,
                  uint32_t n_levels, MeshTreeState* state, uint32_t blocks, void* tmp,
                  hipStream_t s) {
  (void)n_meshes;
  uint64_t* partial = (uint64_t*)tmp;
  // counts[level-1] = elements of this level; counts[level] = elements of the next
  const unsigned long long* n_ptr = counts + (level - 1);
  hipLaunchKernelGGL(k_level_count<K>, dim3(blocks), dim3(BLOCK), 0, s, (const K*)cur.elem,
                     n_ptr, kl, (const MeshTreeState*)state, blocks, partial);
  hipLaunchKernelGGL(k_pair_scan, dim3(1), dim3(BLOCK), 0, s, partial, blocks,
                     (unsigned long long*)nullptr, counts + level);
  hipLaunchKernelGGL(k_level_write<K>, dim3(blocks), dim3(BLOCK), 0, s, level,
                     (const K*)cur.elem, n_ptr, kl, state, blocks, (const uint64_t*)partial,
                     cur.elem_group, (K*)next.elem, next.first_child, spans, n_levels);
}

template void launch_level_first<uint32_t>(const uint32_t*, const unsigned long long*,
                                           KeyLayout<uint32_t>, LevelArrays, LevelArrays,
                                           unsigned long long*, LevelSpan*, uint32_t, uint32_t,
                                           void*, hipStream_t);
template void launch_level_first<uint64_t>(const uint64_t*, const unsigned long long*,
                                           KeyLayout<uint64_t>, LevelArrays, LevelArrays,
                                           unsigned long long*, LevelSpan*, uint32_t, uint32_t,
                                           void*, hipStream_t);
template void launch_level<uint32_t>(int, KeyLayout<uint32_t>, LevelArrays, LevelArrays,
                                     unsigned long long*, LevelSpan*, uint32_t, uint32_t,
                                     MeshTreeState*, uint32_t, void*, hipStream_t);
template void launch_level<uint64_t>(int, KeyLayout<uint64_t>, LevelArrays, LevelArrays,
                                     unsigned long long*, LevelSpan*, uint32_t, uint32_t,
                                     MeshTreeState*, uint32_t, void*, hipStream_t);

}  // namespace svo
