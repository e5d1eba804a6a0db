// k_levels.hip -- sorted unique voxels and the bottom-up octree levels.
//
// Reference: the level loop of SparseVoxelOctree::create (SVO.cu:383-495).
// Per level it flags parent-group heads (cpNumNodes, SVO.cu:221-235), scans
// them (SVO.cu:433-437), places node i at sum[i] + (code & 7) (createNode_*,
// SVO.cu:237-309), fills the empty siblings (createRemainNode, 311-339) and
// stops at the first level holding a single code (487-490).
//
// Here every level is one pass over the level's sorted unique keys:
//   group head  = first key of its parent group (parent differs from the
//                 previous key's parent; composite keys keep meshes apart),
//   root        = the only key of its mesh at this level (the reference's
//                 single-code level), which ends that mesh's tree,
//   output      = next level's keys (the group parents, already sorted and
//                 unique) and, per group, the index of its first child.
// Level 1 is fused with the de-duplication of the sorted hit keys.  Node
// records are written later from these arrays (k_attrs.hip), once the node
// counts are known.
//
// Each pass is a single-pass scan: persistent workgroups claim 4096-element
// tiles by ticket, each thread loads 16 consecutive keys with vector loads,
// the tile's counts are scanned in the workgroup and its offset found by
// decoupled look-back (svo_scan.h); compacted outputs are staged in LDS and
// written contiguously.  The element count is read from device memory, so no
// launch shape depends on it.
#include <algorithm>

#include "svo_internal.h"
#include "svo_scan.h"

namespace svo {
namespace {

constexpr int BLOCK = 256;
constexpr int EPT = 16;             // consecutive elements per thread
constexpr int TILE = BLOCK * EPT;   // 4096
constexpr int MAX_TICKETS = 64;     // one per pass: level 1 .. level 63

// flags of sorted hit key `cur`: bit0 = first copy of its key, bit1 = head of
// its parent group
template <typename K>
__device__ inline uint32_t first_flags_v(K prev, K cur, bool has_prev, KeyLayout<K> kl) {
  if (!has_prev) return 3u;
  if (cur == prev) return 0u;
  return 1u | (kl.parent(cur) != kl.parent(prev) ? 2u : 0u);
}

// flags of level element `cur` of a mesh still being built: bit1 = group
// head, bit2 = the only element of its mesh (the root)
template <typename K>
__device__ inline uint32_t level_flags_v(K prev, K cur, K next, bool has_prev, bool has_next,
                                         KeyLayout<K> kl) {
  const uint32_t m = kl.mesh(cur);
  const bool first = !has_prev || kl.mesh(prev) != m;
  const bool last = !has_next || kl.mesh(next) != m;
  if (first && last) return 4u;
  if (first) return 2u;
  return kl.parent(cur) != kl.parent(prev) ? 2u : 0u;
}

__global__ __launch_bounds__(BLOCK) void k_level_clear(uint64_t* status, uint32_t* tickets,
                                                       const unsigned long long* n_ptr) {
  const uint64_t words = 2 * ((*n_ptr + TILE - 1) / TILE);
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < words; i += stride)
    status[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x < MAX_TICKETS) tickets[threadIdx.x] = 0;
}

// Per-tile frame shared by both passes: claims a tile, loads the thread's run
// and the keys just before and after it.
template <typename K>
struct TileRun {
  uint64_t tile, tb, i0;
  uint32_t tn, cnt;
  K kk[EPT];
  K prev, next;
  bool has_prev, has_next;
};

template <typename K>
__device__ inline bool claim_tile(const K* src, uint64_t n, uint32_t* ticket, uint32_t* s_tile,
                                  K (*s_edge)[BLOCK], TileRun<K>& r) {
  if (threadIdx.x == 0) *s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  r.tile = *s_tile;
  r.tb = r.tile * TILE;
  if (r.tb >= n) return false;  // uniform over the workgroup
  r.tn = (uint32_t)((n - r.tb) < TILE ? (n - r.tb) : TILE);
  const uint32_t l0 = threadIdx.x * EPT;
  r.cnt = l0 < r.tn ? ((r.tn - l0) < EPT ? (r.tn - l0) : EPT) : 0u;
  r.i0 = r.tb + l0;
  load_consecutive(src + r.i0, ((uintptr_t)src & 15) == 0, r.cnt, r.kk);
  K last = r.kk[0];
#pragma unroll
  for (int e = 1; e < EPT; ++e)
    if ((uint32_t)e < r.cnt) last = r.kk[e];
  s_edge[0][threadIdx.x] = r.kk[0];
  s_edge[1][threadIdx.x] = last;
  __syncthreads();
  r.has_prev = r.cnt && r.i0 > 0;
  r.prev = !r.has_prev ? (K)0 : threadIdx.x > 0 ? s_edge[1][threadIdx.x - 1] : src[r.i0 - 1];
  r.has_next = r.cnt && r.i0 + r.cnt < n;
  r.next = !r.has_next ? (K)0
           : (threadIdx.x + 1 < BLOCK && l0 + EPT < r.tn) ? s_edge[0][threadIdx.x + 1]
                                                           : src[r.i0 + r.cnt];
  return true;
}

// ---------------------------------------------------------------- level 1
// Sorted hit keys -> unique keys (elem1) with their group (group1), the level-2
// keys (elem2 = group parents) with their first child (first_child2), the
// per-mesh spans of level 1, and counts[0] / counts[1].
template <typename K>
__global__ __launch_bounds__(BLOCK) void k_first_pass(
    const K* sorted, const unsigned long long* n_ptr, KeyLayout<K> kl, K* elem1,
    uint32_t* group1, K* elem2, uint32_t* first_child2, LevelSpan* spans, uint32_t n_levels,
    uint64_t* status, uint32_t* ticket, uint32_t epoch, unsigned long long* counts) {
  __shared__ K s_edge[2][BLOCK];
  __shared__ K s_key[TILE];
  __shared__ uint32_t s_val[TILE];
  __shared__ uint64_t tmp[4];
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_base[2];
  const uint64_t n = *n_ptr;
  TileRun<K> r;
  while (claim_tile(sorted, n, ticket, &s_tile, s_edge, r)) {
    uint32_t f = 0, nu = 0, ng = 0;
    {
      K pk = r.prev;
      bool hp = r.has_prev;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        if ((uint32_t)e < r.cnt) {
          const uint32_t fl = first_flags_v(pk, r.kk[e], hp, kl);
          f |= fl << (2 * e);
          nu += fl & 1u;
          ng += fl >> 1;
          pk = r.kk[e];
          hp = true;
        }
      }
    }
    uint64_t tot;
    const uint64_t ex = block_excl_scan256(((uint64_t)ng << 32) | nu, tmp, tot);
    const uint32_t tu = (uint32_t)tot, tg = (uint32_t)(tot >> 32);
    if (threadIdx.x < 2) {
      const uint32_t c = threadIdx.x == 0 ? tu : tg;
      const uint32_t pre = lookback_exclusive(status + threadIdx.x, 2, r.tile, c, epoch);
      s_base[threadIdx.x] = pre;
      if (r.tb + r.tn == n) counts[threadIdx.x] = (unsigned long long)pre + c;
    }
    __syncthreads();
    const uint32_t bu = s_base[0], bg = s_base[1];
    // elem1 / group1 staged by tile-local unique index; spans on the way
    {
      uint32_t ru = (uint32_t)ex, rg = (uint32_t)(ex >> 32);
      K pk = r.prev;
      bool hp = r.has_prev;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        if ((uint32_t)e < r.cnt) {
          const K key = r.kk[e];
          const uint32_t fl = (f >> (2 * e)) & 3u;
          const uint32_t m = kl.mesh(key);
          if (fl & 1u) {
            if (fl & 2u) ++rg;
            s_key[ru] = key;
            s_val[ru] = bg + rg - 1;
            if (!hp || kl.mesh(pk) != m) {  // first element of its mesh
              spans[(uint64_t)m * n_levels].elem_first = bu + ru;
              spans[(uint64_t)m * n_levels].group_first = bg + rg - 1;
            }
            ++ru;
          }
          const bool hn = (uint32_t)e + 1 < r.cnt || r.has_next;
          const K nk = (uint32_t)e + 1 < r.cnt ? r.kk[e + 1 < EPT ? e + 1 : e] : r.next;
          if (!hn || kl.mesh(nk) != m) {  // last element (copy) of its mesh
            spans[(uint64_t)m * n_levels].elem_end = bu + ru;
            spans[(uint64_t)m * n_levels].group_end = bg + rg;
          }
          pk = key;
          hp = true;
        }
      }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < tu; i += BLOCK) {
      elem1[(uint64_t)bu + i] = s_key[i];
      group1[(uint64_t)bu + i] = s_val[i];
    }
    __syncthreads();
    // elem2 / first_child2 staged by tile-local group index
    {
      uint32_t ru = (uint32_t)ex, rg = (uint32_t)(ex >> 32);
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        if ((uint32_t)e < r.cnt) {
          const uint32_t fl = (f >> (2 * e)) & 3u;
          if (fl & 1u) {
            if (fl & 2u) {
              s_key[rg] = kl.parent(r.kk[e]);
              s_val[rg] = bu + ru;
              ++rg;
            }
            ++ru;
          }
        }
      }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < tg; i += BLOCK) {
      elem2[(uint64_t)bg + i] = s_key[i];
      first_child2[(uint64_t)bg + i] = s_val[i];
    }
    __syncthreads();
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
// Level elements -> their group (group_l, NONE for roots and finished
// meshes), the next level's keys with their first child, the spans of this
// level, roots into state, and counts[level] (elements of the next level).
template <typename K>
__global__ __launch_bounds__(BLOCK) void k_level_pass(
    int level, const K* elem, const unsigned long long* n_ptr, KeyLayout<K> kl,
    MeshTreeState* state, uint32_t* group_l, K* elem_next, uint32_t* first_child_next,
    LevelSpan* spans, uint32_t n_levels, uint64_t* status, uint32_t* ticket, uint32_t epoch,
    unsigned long long* count_next) {
  __shared__ K s_edge[2][BLOCK];
  __shared__ K s_key[TILE];
  __shared__ uint32_t s_val[TILE];
  __shared__ uint32_t tmp[4];
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_base;
  const uint64_t n = *n_ptr;
  const uint32_t li = (uint32_t)(level - 1);
  const bool vec_out = ((uintptr_t)group_l & 15) == 0;
  TileRun<K> r;
  while (claim_tile(elem, n, ticket, &s_tile, s_edge, r)) {
    // 4 flag bits per element: bit0 live (its mesh is still being built),
    // bit1 group head, bit2 root
    uint64_t f = 0;
    uint32_t ng = 0;
    {
      uint32_t last_m = 0xFFFFFFFFu, done = 0;
      K pk = r.prev;
      bool hp = r.has_prev;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        if ((uint32_t)e < r.cnt) {
          const K cur = r.kk[e];
          const uint32_t m = kl.mesh(cur);
          if (m != last_m) {
            done = state[m].depth;
            last_m = m;
          }
          if (done == 0) {
            const bool more = e + 1 < EPT && (uint32_t)(e + 1) < r.cnt;
            const bool hn = more || r.has_next;
            const K nk = more ? r.kk[(e + 1) % EPT] : r.next;
            const uint32_t fl = level_flags_v(pk, cur, nk, hp, hn, kl) | 1u;
            f |= (uint64_t)fl << (4 * e);
            ng += (fl >> 1) & 1u;
          }
          pk = cur;
          hp = true;
        }
      }
    }
    uint32_t tg;
    const uint32_t exg = block_excl_scan256(ng, tmp, tg);
    if (threadIdx.x == 0) {
      const uint32_t pre = lookback_exclusive(status + 1, 2, r.tile, tg, epoch);
      s_base = pre;
      if (r.tb + r.tn == n) *count_next = (unsigned long long)pre + tg;
    }
    __syncthreads();
    const uint32_t bg = s_base;
    uint32_t gv[EPT];
    {
      uint32_t rg = exg;
      K pk = r.prev;
      bool hp = r.has_prev;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        gv[e] = NONE;
        if ((uint32_t)e < r.cnt) {
          const K key = r.kk[e];
          const uint64_t i = r.i0 + e;
          const uint32_t fl = (uint32_t)(f >> (4 * e)) & 15u;
          const uint32_t m = kl.mesh(key);
          if (fl & 4u) {  // the only key of its mesh at this level: the root
            state[m].depth = (uint32_t)level;
            state[m].root_elem = (uint32_t)i;
            LevelSpan& sp = spans[(uint64_t)m * n_levels + li];
            sp.elem_first = (uint32_t)i;
            sp.elem_end = (uint32_t)i + 1;
            sp.group_first = sp.group_end = 0;
          } else if (fl & 1u) {
            if (fl & 2u) {
              s_key[rg] = kl.parent(key);
              s_val[rg] = (uint32_t)i;
              ++rg;
            }
            gv[e] = bg + rg - 1;
            const bool more = e + 1 < EPT && (uint32_t)(e + 1) < r.cnt;
            const bool hn = more || r.has_next;
            const K nk = more ? r.kk[(e + 1) % EPT] : r.next;
            LevelSpan& sp = spans[(uint64_t)m * n_levels + li];
            if (!hp || kl.mesh(pk) != m) {
              sp.elem_first = (uint32_t)i;
              sp.group_first = bg + rg - 1;
            }
            if (!hn || kl.mesh(nk) != m) {
              sp.elem_end = (uint32_t)i + 1;
              sp.group_end = bg + rg;
            }
          }
          pk = key;
          hp = true;
        }
      }
    }
    if (vec_out && r.cnt == EPT) {
      uint4* q = (uint4*)(group_l + r.i0);
#pragma unroll
      for (int v = 0; v < EPT / 4; ++v)
        q[v] = make_uint4(gv[4 * v], gv[4 * v + 1], gv[4 * v + 2], gv[4 * v + 3]);
    } else {
#pragma unroll
      for (int e = 0; e < EPT; ++e)
        if ((uint32_t)e < r.cnt) group_l[r.i0 + e] = gv[e];
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < tg; i += BLOCK) {
      elem_next[(uint64_t)bg + i] = s_key[i];
      first_child_next[(uint64_t)bg + i] = s_val[i];
    }
    __syncthreads();
  }
}

uint32_t grid_tiles(uint64_t cap, uint32_t blocks) {
  const uint64_t tiles = (cap + TILE - 1) / TILE;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(tiles, blocks));
}

}  // namespace

size_t level_tmp_bytes(uint64_t cap) {
  const uint64_t tiles = (std::max<uint64_t>(cap, 1) + TILE - 1) / TILE;
  return 256 + 2 * tiles * sizeof(uint64_t);
}

template <typename K>
void launch_level_first(const K* sorted, const unsigned long long* n_sorted, uint64_t cap,
                        KeyLayout<K> kl, LevelArrays l1, LevelArrays l2,
                        unsigned long long* counts, LevelSpan* spans1, uint32_t n_levels,
                        uint32_t blocks, void* tmp, hipStream_t s) {
  uint32_t* tickets = (uint32_t*)tmp;
  uint64_t* status = (uint64_t*)((char*)tmp + 256);
  const uint64_t tiles = (std::max<uint64_t>(cap, 1) + TILE - 1) / TILE;
  const uint32_t cg = (uint32_t)std::min<uint64_t>(2048, (2 * tiles + BLOCK - 1) / BLOCK);
  hipLaunchKernelGGL(k_level_clear, dim3(std::max<uint32_t>(cg, 1)), dim3(BLOCK), 0, s, status,
                     tickets, n_sorted);
  hipLaunchKernelGGL(k_first_pass<K>, dim3(grid_tiles(cap, blocks)), dim3(BLOCK), 0, s, sorted,
                     n_sorted, kl, (K*)l1.elem, l1.elem_group, (K*)l2.elem, l2.first_child,
                     spans1, n_levels, status, tickets, 1u, counts);
}

void launch_level_first_finalize(const LevelSpan* spans, uint32_t n_meshes, uint32_t n_levels,
                                 MeshTreeState* state, hipStream_t s) {
  hipLaunchKernelGGL(k_first_finalize, dim3((n_meshes + 63) / 64), dim3(64), 0, s, spans,
                     n_meshes, n_levels, state);
}

template <typename K>
void launch_level(int level, KeyLayout<K> kl, LevelArrays cur, LevelArrays next,
                  unsigned long long* counts, LevelSpan* spans, uint32_t n_levels,
                  MeshTreeState* state, uint32_t blocks, void* tmp, hipStream_t s) {
  uint32_t* tickets = (uint32_t*)tmp;
  uint64_t* status = (uint64_t*)((char*)tmp + 256);
  // counts[level-1] = elements of this level; counts[level] = elements of the next
  hipLaunchKernelGGL(k_level_pass<K>, dim3(grid_tiles(cur.cap, blocks)), dim3(BLOCK), 0, s,
                     level, (const K*)cur.elem, counts + (level - 1), kl, state,
                     cur.elem_group, (K*)next.elem, next.first_child, spans, n_levels, status,
                     tickets + (level - 1), (uint32_t)level, counts + level);
}

template void launch_level_first<uint32_t>(const uint32_t*, const unsigned long long*, uint64_t,
                                           KeyLayout<uint32_t>, LevelArrays, LevelArrays,
                                           unsigned long long*, LevelSpan*, uint32_t, uint32_t,
                                           void*, hipStream_t);
template void launch_level_first<uint64_t>(const uint64_t*, const unsigned long long*, uint64_t,
                                           KeyLayout<uint64_t>, LevelArrays, LevelArrays,
                                           unsigned long long*, LevelSpan*, uint32_t, uint32_t,
                                           void*, hipStream_t);
template void launch_level<uint32_t>(int, KeyLayout<uint32_t>, LevelArrays, LevelArrays,
                                     unsigned long long*, LevelSpan*, uint32_t, MeshTreeState*,
                                     uint32_t, void*, hipStream_t);
template void launch_level<uint64_t>(int, KeyLayout<uint64_t>, LevelArrays, LevelArrays,
                                     unsigned long long*, LevelSpan*, uint32_t, MeshTreeState*,
                                     uint32_t, void*, hipStream_t);

}  // namespace svo
