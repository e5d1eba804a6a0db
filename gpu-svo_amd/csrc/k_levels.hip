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
// Sorted hit keys -> unique keys (elem1), the level-2 keys (elem2 = group
// parents) with their first child (first_child2), the per-mesh spans of
// level 1, and counts[0] / counts[1].  (Level 1's own group index per key is
// not needed: node records look groups up from the level above.)
template <typename K>
__global__ __launch_bounds__(BLOCK) void k_first_pass(
    const K* sorted, const unsigned long long* n_ptr, KeyLayout<K> kl, K* elem1, K* elem2,
    uint32_t* first_child2, LevelSpan* spans, uint32_t n_levels,
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
    if (threadIdx.x < 128) {
      // wave 0: unique-key offset, wave 1: group offset
      const uint32_t wv = threadIdx.x >> 6;
      const uint32_t c = wv == 0 ? tu : tg;
      const uint32_t pre = lookback_exclusive_wave(status + wv, 2, r.tile, c, epoch);
      if (__lane_id() == 0) {
        s_base[wv] = pre;
        if (r.tb + r.tn == n) counts[wv] = (unsigned long long)pre + c;
      }
    }
    __syncthreads();
    const uint32_t bu = s_base[0], bg = s_base[1];
    // elem1 staged by tile-local unique index; spans on the way
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
    for (uint32_t i = threadIdx.x; i < tu; i += BLOCK) elem1[(uint64_t)bu + i] = s_key[i];
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

// ---------------------------------------------------------------- level 1 from the bitmap
// The occupancy bitmap holds, per mesh, one bit per cell in Morton order, so
// its set bits in word order are the sorted unique keys and the nonzero bytes
// are the parent groups (the eight children of parent p are codes 8p..8p+7,
// one byte).  count -> scan of tile counts -> emit, each over tiles of 2048
// words.
constexpr int BM_WPT = 8;                 // words per thread (two 16-byte pieces)
constexpr int BM_TILE = BLOCK * BM_WPT;   // 2048 words
static_assert(BM_WPT % 4 == 0, "threads load whole 16-byte pieces");

__device__ inline uint32_t nonzero_bytes(uint32_t w) {
  const uint32_t t = (w & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;  // bit 7 of a byte: low 7 bits nonzero
  return __popc((t | w) & 0x80808080u);
}

__device__ inline void load_words(const uint32_t* bits, uint64_t n_words, uint64_t w0,
                                  uint32_t (&w)[BM_WPT]) {
  if (w0 + BM_WPT <= n_words) {
#pragma unroll
    for (int q = 0; q < BM_WPT / 4; ++q) {
      const uint4 a = ((const uint4*)(bits + w0))[q];
      w[4 * q] = a.x; w[4 * q + 1] = a.y; w[4 * q + 2] = a.z; w[4 * q + 3] = a.w;
    }
  } else {
#pragma unroll
    for (int j = 0; j < BM_WPT; ++j) w[j] = w0 + j < n_words ? bits[w0 + j] : 0u;
  }
}

// Zeroes the 16-byte pieces of the words loaded by load_words that are not
// already zero.
__device__ inline void clear_words(uint32_t* bits, uint64_t n_words, uint64_t w0,
                                   const uint32_t (&w)[BM_WPT]) {
  if (w0 + BM_WPT <= n_words) {
#pragma unroll
    for (int q = 0; q < BM_WPT / 4; ++q)
      if (w[4 * q] | w[4 * q + 1] | w[4 * q + 2] | w[4 * q + 3])
        ((uint4*)(bits + w0))[q] = make_uint4(0u, 0u, 0u, 0u);
  } else {
#pragma unroll
    for (int j = 0; j < BM_WPT; ++j)
      if (w0 + j < n_words && w[j]) bits[w0 + j] = 0u;
  }
}

// Tile counts per thread in the scan: short chunks spread the loads over
// more workgroups (4 measured faster than 16, and 64, one workgroup for a
// res-1024 bitmap, slower still: its loads are latency-bound).
constexpr int SC_PER = 4;
constexpr int SC_CHUNK = BLOCK * SC_PER;     // 1024 tiles per scan workgroup

// Per tile: set bits, nonzero bytes, and the 4-byte units k_bm_emit will
// zero.
__global__ __launch_bounds__(BLOCK) void k_bm_count(const uint32_t* bits, uint64_t n_words,
                                                    uint32_t* tile_u, uint32_t* tile_g,
                                                    uint32_t* tile_r, uint64_t* scan_status,
                                                    uint32_t n_status) {
  __shared__ uint32_t tmp[8];
  // the tile-count scan's look-back words and ticket start at zero
  if (blockIdx.x == 0)
    for (uint32_t i = threadIdx.x; i < n_status; i += BLOCK) scan_status[i] = 0;
  const uint64_t w0 = (uint64_t)blockIdx.x * BM_TILE + threadIdx.x * BM_WPT;
  uint32_t w[BM_WPT];
  load_words(bits, n_words, w0, w);
  // a: set bits (0-16) and the 4-byte units clear_words will zero (17-28);
  // b: nonzero bytes
  uint32_t a = 0, b = 0;
#pragma unroll
  for (int j = 0; j < BM_WPT; ++j) {
    a += __popc(w[j]);
    b += nonzero_bytes(w[j]);
  }
  if (w0 + BM_WPT <= n_words) {
#pragma unroll
    for (int q = 0; q < BM_WPT / 4; ++q)
      a += (w[4 * q] | w[4 * q + 1] | w[4 * q + 2] | w[4 * q + 3]) ? (4u << 17) : 0u;
  } else {
#pragma unroll
    for (int j = 0; j < BM_WPT; ++j) a += (uint32_t)(w[j] != 0) << 17;
  }
  uint32_t ta, tb;
  block_excl_scan256_2(a, b, tmp, ta, tb);  // only the totals are used
  if (threadIdx.x == 0) {
    tile_u[blockIdx.x] = ta & 0x1FFFFu;
    tile_g[blockIdx.x] = tb;
    tile_r[blockIdx.x] = ta >> 17;
  }
}

// Exclusive scan of the tile counts in place, SC_CHUNK tiles per workgroup, chunk
// offsets by decoupled look-back (chunks claimed by ticket); totals to
// counts[0] (unique keys = level-1 elements) and counts[1] (groups = level-2
// elements); the bytes k_bm_emit will zero are added to *reset_bytes.
__global__ __launch_bounds__(BLOCK) void k_bm_scan(uint32_t* tile_u, uint32_t* tile_g,
                                                   const uint32_t* tile_r, uint32_t tiles,
                                                   uint64_t* scan_status,
                                                   unsigned long long* counts,
                                                   unsigned long long* reset_bytes) {
  __shared__ uint64_t tmp[4];
  __shared__ uint32_t s_chunk, s_base[2], s_reset[BLOCK / 64];
  uint32_t* ticket = (uint32_t*)scan_status;  // word 0; chunk c uses words 2 + 2c, 3 + 2c
  if (threadIdx.x == 0) s_chunk = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t chunk = s_chunk;
  const uint32_t i0 = chunk * SC_CHUNK + threadIdx.x * SC_PER;
  uint32_t a[SC_PER], b[SC_PER];
#pragma unroll
  for (int j = 0; j < SC_PER; ++j) {
    const uint32_t i = i0 + j;
    const uint32_t c = i < tiles ? i : tiles - 1;
    a[j] = tile_u[c];
    b[j] = tile_g[c];
  }
  uint32_t su = 0, sg = 0, sr = 0;
#pragma unroll
  for (int j = 0; j < SC_PER; ++j)
    if (i0 + j < tiles) {
      su += a[j];
      sg += b[j];
      sr += tile_r[i0 + j];
    }
  for (int d = 32; d >= 1; d >>= 1) sr += __shfl_xor(sr, d);
  if (__lane_id() == 0) s_reset[threadIdx.x >> 6] = sr;
  uint64_t tot;
  const uint64_t ex = block_excl_scan256(((uint64_t)sg << 32) | su, tmp, tot);
  if (threadIdx.x < 2) {
    const uint32_t c = threadIdx.x == 0 ? (uint32_t)tot : (uint32_t)(tot >> 32);
    const uint32_t pre = lookback_exclusive(scan_status + 2 + threadIdx.x, 2, chunk, c, 1u);
    s_base[threadIdx.x] = pre;
    if ((uint64_t)(chunk + 1) * SC_CHUNK >= tiles) counts[threadIdx.x] = (unsigned long long)pre + c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t r = 0;
    for (int v = 0; v < BLOCK / 64; ++v) r += s_reset[v];
    if (r) atomicAdd(reset_bytes, (unsigned long long)(4 * r));
  }
  uint32_t ru = s_base[0] + (uint32_t)ex, rg = s_base[1] + (uint32_t)(ex >> 32);
#pragma unroll
  for (int j = 0; j < SC_PER; ++j) {
    const uint32_t i = i0 + j;
    if (i < tiles) {
      tile_u[i] = ru;
      tile_g[i] = rg;
      ru += a[j];
      rg += b[j];
    }
  }
}

__device__ inline uint32_t mesh_of_word(const uint64_t* bm_off, uint32_t n_meshes, uint64_t gw) {
  uint32_t lo = 0, hi = n_meshes;  // bm_off[lo] <= gw < bm_off[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (bm_off[mid] <= gw) lo = mid; else hi = mid;
  }
  return lo;
}

// One tile of 2048 words, which lies inside one mesh's bitmap region (the
// regions start on tile boundaries).  The tile's nonzero words are compacted
// into LDS in word order with the exclusive prefixes of their set bits (low
// 17 bits) and nonzero bytes (high 15 bits).  Then, 256 nonzero words per
// round, a lane takes one word and walks its set bits (unique keys, in order)
// and its nonzero bytes (groups) into LDS buffers, which the block writes out
// contiguously.  A lane per nonzero word keeps most lanes busy (a lane per
// bit would leave about four in five idle at surface densities), and staging
// turns the lanes' strided stores into whole cache lines.  A key is staged
// as its 16-bit bit index in the tile (the tile's first key is added at
// copy-out) and a group as the index of its first key only (its parent is
// read back from that key), so the workgroup's LDS stays near 26 KB for
// either key width and six workgroups fit a CU.  A round with more than
// EMIT_CAP keys (dense tiles) stores its keys and groups directly.  The
// tile's words are zeroed after they are read.
constexpr uint32_t EMIT_CAP = 2048;
static_assert((EMIT_CAP & (EMIT_CAP - 1)) == 0, "the copy-out masks slot indices");
static_assert(BM_TILE == BITMAP_TILE_WORDS, "mesh regions start on tile boundaries");

template <typename K>
__global__ __launch_bounds__(BLOCK) void k_bm_emit(uint32_t* bits, uint64_t n_words,
                                                   const uint64_t* bm_off, uint32_t n_meshes,
                                                   KeyLayout<K> kl, const uint32_t* tile_u,
                                                   const uint32_t* tile_g, K* elem1, K* elem2,
                                                   uint32_t* first_child2,
                                                   const unsigned long long* counts,
                                                   LevelSpan* spans, uint32_t n_levels,
                                                   MeshTreeState* state, uint64_t* status,
                                                   uint32_t* tickets) {
  __shared__ uint32_t tmp[8];
  __shared__ uint32_t s_word[BM_TILE], s_pre[BM_TILE];
  __shared__ uint16_t s_idx[BM_TILE], s_key[EMIT_CAP], s_fc[BLOCK * 4];
  __shared__ K s_kbase;  // key of the tile's bit 0
  const uint64_t tw0 = (uint64_t)blockIdx.x * BM_TILE;
  const uint32_t l0 = threadIdx.x * BM_WPT;
  // the bitmap words and the tile's output offsets are loaded first, so that
  // the mesh lookup's dependent loads run while they are in flight
  uint32_t w[BM_WPT];
  load_words(bits, n_words, tw0 + l0, w);
  const uint32_t bu = tile_u[blockIdx.x], bg = tile_g[blockIdx.x];
  if (threadIdx.x == 0) {
    const uint32_t m = mesh_of_word(bm_off, n_meshes, tw0);
    const uint64_t off = bm_off[m];
    s_kbase = ((K)m << kl.cb) | ((K)(tw0 - off) << 5);
    if (tw0 == off) {
      // the mesh's first tile: its level-1 span ends where the next mesh's
      // first tile starts (every region is whole tiles) and, for a mesh whose
      // level 1 holds a single voxel, its depth-1 root (what
      // k_first_finalize does on the sort path)
      const uint64_t tn = bm_off[m + 1] / BM_TILE;
      const uint32_t ue = tn < gridDim.x ? tile_u[tn] : (uint32_t)counts[0];
      const uint32_t ge = tn < gridDim.x ? tile_g[tn] : (uint32_t)counts[1];
      LevelSpan& sp = spans[(uint64_t)m * n_levels];
      sp.elem_first = bu;
      sp.elem_end = ue;
      sp.group_first = bg;
      sp.group_end = ge;
      if (ue - bu == 1) {
        state[m].depth = 1;
        state[m].root_elem = bu;
      }
    }
  }
  {
    // the later levels' look-back words and tickets start at zero (as
    // k_level_clear leaves them), a slice per workgroup
    const uint64_t words = 2 * ((counts[0] + TILE - 1) / TILE);
    const uint64_t per = (words + gridDim.x - 1) / gridDim.x;
    const uint64_t i1 = std::min<uint64_t>(words, (uint64_t)(blockIdx.x + 1) * per);
    for (uint64_t i = (uint64_t)blockIdx.x * per + threadIdx.x; i < i1; i += BLOCK) status[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x < MAX_TICKETS) tickets[threadIdx.x] = 0;
  }
  clear_words(bits, n_words, tw0 + l0, w);  // the bitmap is left clear for the next pass
  // a: set bits (0-16) and nonzero words (17-28); b: nonzero bytes
  uint32_t a = 0, b = 0;
#pragma unroll
  for (int j = 0; j < BM_WPT; ++j) {
    a += __popc(w[j]) + ((uint32_t)(w[j] != 0) << 17);
    b += nonzero_bytes(w[j]);
  }
  uint32_t ta, tb;
  block_excl_scan256_2(a, b, tmp, ta, tb);  // tmp is not used again
  {
    uint32_t pu = a & 0x1FFFFu, pg = b, pz = a >> 17;
#pragma unroll
    for (int j = 0; j < BM_WPT; ++j) {
      if (w[j]) {
        s_word[pz] = w[j];
        s_pre[pz] = pu | (pg << 17);  // pu <= 65535, pg <= 8191 for a nonzero word
        s_idx[pz] = (uint16_t)(l0 + j);
        ++pz;
      }
      pu += __popc(w[j]);
      pg += nonzero_bytes(w[j]);
    }
  }
  __syncthreads();
  const uint32_t nz = ta >> 17;
  const uint32_t pre_end = (ta & 0x1FFFFu) | (tb << 17);
  const K kbase = s_kbase;
  for (uint32_t r0 = 0; r0 < nz; r0 += BLOCK) {
    const uint32_t r1 = r0 + BLOCK < nz ? r0 + BLOCK : nz;
    const uint32_t pre0 = s_pre[r0], pre1 = r1 < nz ? s_pre[r1] : pre_end;
    const uint32_t u0 = pre0 & 0x1FFFFu, nu = (pre1 & 0x1FFFFu) - u0;
    const uint32_t g0 = pre0 >> 17, ng = (pre1 >> 17) - g0;
    const bool stage = nu <= EMIT_CAP;  // the same for the whole block
    const uint32_t it = r0 + threadIdx.x;
    if (it < r1) {
      const uint32_t word = s_word[it], pre = s_pre[it];
      const uint32_t wb = (uint32_t)s_idx[it] << 5;  // tile bit index of the word's bit 0
      const uint32_t lu = (pre & 0x1FFFFu) - u0;
      // unique keys: the set bits in order
      if (stage) {
        uint32_t o = lu;
        for (uint32_t r = word; r; r &= r - 1u) s_key[o++] = (uint16_t)(wb | __builtin_ctz(r));
      } else {
        K* out = elem1 + (uint64_t)bu + u0 + lu;
        const K hi = kbase + (K)wb;
        for (uint32_t r = word; r; r &= r - 1u) *out++ = hi | (K)__builtin_ctz(r);
      }
      // groups: the nonzero bytes in order, tested in turn (a loop over the
      // nonzero bytes runs as often as the wave's fullest word has them,
      // nearly always four).  A group's first child is the unique key of
      // the byte's lowest set bit (the set bits below it are those of the
      // lower bytes); its parent is parent(word's key 0) + byte index.
      uint32_t lg = (pre >> 17) - g0;
      const uint32_t pu_w = pre & 0x1FFFFu;
      if (stage) {
#pragma unroll
        for (uint32_t b = 0; b < 4; ++b) {
          if ((word >> (8 * b)) & 0xFFu) {
            s_fc[lg] = (uint16_t)(pu_w + (b ? __popc(word & ((1u << (8 * b)) - 1u)) : 0u));
            ++lg;
          }
        }
      } else {
        const K phi = kl.parent(kbase + (K)wb);
        K* dp = elem2 + (uint64_t)bg + g0;
        uint32_t* df = first_child2 + (uint64_t)bg + g0;
#pragma unroll
        for (uint32_t b = 0; b < 4; ++b) {
          if ((word >> (8 * b)) & 0xFFu) {
            dp[lg] = phi | (K)b;
            df[lg] = bu + pu_w + (b ? __popc(word & ((1u << (8 * b)) - 1u)) : 0u);
            ++lg;
          }
        }
      }
    }
    __syncthreads();
    // copy-out, four elements per thread per step: the LDS reads are issued
    // together (slots past the round's count read stale entries inside the
    // buffers and store nothing)
    if (stage) {
      K* dst = elem1 + (uint64_t)bu + u0;
      for (uint32_t i0 = threadIdx.x; i0 < nu; i0 += 4 * BLOCK) {
        uint32_t v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = s_key[(i0 + j * BLOCK) & (EMIT_CAP - 1)];
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (i0 + j * BLOCK < nu) dst[i0 + j * BLOCK] = kbase + (K)v[j];
      }
      // groups, ng <= 4 * BLOCK: one step
      const uint32_t i0 = threadIdx.x;
      uint32_t f[4], v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) f[j] = s_fc[i0 + j * BLOCK];
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] = s_key[(f[j] - u0) & (EMIT_CAP - 1)];
      K* dp = elem2 + (uint64_t)bg + g0;
      uint32_t* df = first_child2 + (uint64_t)bg + g0;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (i0 + j * BLOCK < ng) {
          dp[i0 + j * BLOCK] = kl.parent(kbase + (K)v[j]);
          df[i0 + j * BLOCK] = bu + f[j];
        }
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
    if (threadIdx.x < 64) {
      const uint32_t pre = lookback_exclusive_wave(status + 1, 2, r.tile, tg, epoch);
      if (threadIdx.x == 0) {
        s_base = pre;
        if (r.tb + r.tn == n) *count_next = (unsigned long long)pre + tg;
      }
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
                     n_sorted, kl, (K*)l1.elem, (K*)l2.elem, l2.first_child,
                     spans1, n_levels, status, tickets, 1u, counts);
}

size_t bitmap_level_tmp_bytes(uint64_t n_words) {
  const uint64_t tiles = (std::max<uint64_t>(n_words, 1) + BM_TILE - 1) / BM_TILE;
  const uint64_t chunks = (tiles + SC_CHUNK - 1) / SC_CHUNK;
  return 3 * tiles * sizeof(uint32_t) + 256 + (2 + 2 * chunks) * sizeof(uint64_t);
}

template <typename K>
void launch_level_first_bitmap(uint32_t* bits, uint64_t n_words, const uint64_t* bm_off,
                               uint32_t n_meshes, KeyLayout<K> kl, LevelArrays l1,
                               LevelArrays l2, unsigned long long* counts, LevelSpan* spans1,
                               uint32_t n_levels, MeshTreeState* state, void* tmp,
                               void* lvl_tmp, unsigned long long* reset_bytes, hipStream_t s) {
  const uint64_t tiles = (std::max<uint64_t>(n_words, 1) + BM_TILE - 1) / BM_TILE;
  const uint64_t chunks = (tiles + SC_CHUNK - 1) / SC_CHUNK;
  uint32_t* tile_u = (uint32_t*)tmp;
  uint32_t* tile_g = tile_u + tiles;
  uint32_t* tile_r = tile_g + tiles;
  uint64_t* scan_status =
      (uint64_t*)((char*)tmp + ((3 * tiles * sizeof(uint32_t) + 255) / 256) * 256);
  const uint32_t n_status = (uint32_t)(2 + 2 * chunks);
  hipLaunchKernelGGL(k_bm_count, dim3((uint32_t)tiles), dim3(BLOCK), 0, s, bits, n_words, tile_u,
                     tile_g, tile_r, scan_status, n_status);
  hipLaunchKernelGGL(k_bm_scan, dim3((uint32_t)chunks), dim3(BLOCK), 0, s, tile_u, tile_g,
                     (const uint32_t*)tile_r, (uint32_t)tiles, scan_status, counts, reset_bytes);
  // the emission also writes the level-1 spans and depth-1 roots, and clears
  // the later levels' look-back words and tickets (launch_level)
  uint32_t* tickets = (uint32_t*)lvl_tmp;
  uint64_t* status = (uint64_t*)((char*)lvl_tmp + 256);
  hipLaunchKernelGGL(k_bm_emit<K>, dim3((uint32_t)tiles), dim3(BLOCK), 0, s, bits, n_words,
                     bm_off, n_meshes, kl, (const uint32_t*)tile_u, (const uint32_t*)tile_g,
                     (K*)l1.elem, (K*)l2.elem, l2.first_child,
                     (const unsigned long long*)counts, spans1, n_levels, state, status,
                     tickets);
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
template void launch_level_first_bitmap<uint32_t>(uint32_t*, uint64_t, const uint64_t*,
                                                  uint32_t, KeyLayout<uint32_t>, LevelArrays,
                                                  LevelArrays, unsigned long long*, LevelSpan*,
                                                  uint32_t, MeshTreeState*, void*, void*,
                                                  unsigned long long*, hipStream_t);
template void launch_level_first_bitmap<uint64_t>(uint32_t*, uint64_t, const uint64_t*,
                                                  uint32_t, KeyLayout<uint64_t>, LevelArrays,
                                                  LevelArrays, unsigned long long*, LevelSpan*,
                                                  uint32_t, MeshTreeState*, void*, void*,
                                                  unsigned long long*, hipStream_t);
template void launch_level<uint32_t>(int, KeyLayout<uint32_t>, LevelArrays, LevelArrays,
                                     unsigned long long*, LevelSpan*, uint32_t, MeshTreeState*,
                                     uint32_t, void*, hipStream_t);
template void launch_level<uint64_t>(int, KeyLayout<uint64_t>, LevelArrays, LevelArrays,
                                     unsigned long long*, LevelSpan*, uint32_t, MeshTreeState*,
                                     uint32_t, void*, hipStream_t);

}  // namespace svo
