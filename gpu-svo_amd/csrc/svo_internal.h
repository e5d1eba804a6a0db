// svo_internal.h -- host-side declarations shared by the pipeline and the
// kernel translation units.  Every launcher takes device pointers and a
// stream and enqueues work only; none of them synchronises.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "svo_device.h"

namespace svo {

// ------------------------------------------------------------ tuning knobs
constexpr int SIZE_CLASSES = 15;

struct Tuning {
  uint32_t small_max_cells = 1024;  // bbox cells handled by one lane
  uint32_t sort_blocks = 65536;     // grid bound of the radix sort passes (tiles stride)
  uint32_t scan_blocks = 1024;      // fixed grid of the level / attribute scans
  // Occupancy bitmap instead of the sort (level 1 straight from the bitmap)
  // while the bitmap is at most this many bytes and at most this many bytes
  // per candidate cell (so sparse hits in a huge grid still take the sort).
  uint64_t bitmap_max_bytes = 32ull << 30;
  uint32_t bitmap_bytes_per_cell = 64;
  int force_path = 0;               // 0 automatic, 1 bitmap, 2 sort
  // svo_build_batch runs batches of at least 2 * batch_min_meshes meshes as
  // up to batch_lanes concurrent passes
  uint32_t batch_lanes = 2;
  uint32_t batch_min_meshes = 16;
  // Per size class (size_class_of) box-cell limit of the one-lane
  // voxelization; 0 = small_max_cells.  Set by the tuning search.
  uint32_t class_small_max[SIZE_CLASSES] = {};
};

// Size class of a mesh for tuning: 5 resolution bands (max axis <= 64, 128,
// 256, 512, above) x 3 bands of res^2 / triangles, which grows with the
// triangles' edge length in voxels for a surface spanning its box (<= 4:
// triangles of about a voxel or less, <= 64: a few voxels, above: larger).
inline int size_class_of(const int32_t res[3], uint64_t n_tris) {
  const int32_t r = res[0] > res[1] ? (res[0] > res[2] ? res[0] : res[2])
                                    : (res[1] > res[2] ? res[1] : res[2]);
  const int band = r <= 64 ? 0 : r <= 128 ? 1 : r <= 256 ? 2 : r <= 512 ? 3 : 4;
  const double q = (double)r * (double)r / (double)(n_tris ? n_tris : 1);
  const int qb = q <= 4.0 ? 0 : q <= 64.0 ? 1 : 2;
  return 3 * band + qb;
}

// ------------------------------------------------------------ k_voxelize.hip
// Per-mesh ordered-float min/max of all points (6 x uint32 per mesh) and the
// grid parameters derived from it (BaseModel.h:24-73, SVO.cu:342-351).
void launch_bbox(const MeshDev* meshes, uint32_t n_meshes, uint64_t max_points, uint32_t* acc,
                 hipStream_t s);
void launch_params(const MeshDev* meshes, uint32_t n_meshes, const uint32_t* acc,
                   MeshParams* params, hipStream_t s);
// Candidate cells and big-triangle chunk counts (2 x uint64 added to
// `totals`), and the order launch_voxelize takes the triangles in
// (`order`: n_tris uint32, a permutation of the triangle indices that
// groups triangles of about one box size).  `order` heads a buffer of
// order_bytes(n_tris) bytes whose tail `bins` (at byte 4 * n_tris rounded up
// to 256) holds the triangles' size classes.
size_t order_bytes(uint64_t n_tris);
void launch_classify(const MeshDev* meshes, const MeshParams* params, const uint64_t* tri_base,
                     uint32_t n_meshes, uint64_t n_tris, const Tuning& tn,
                     unsigned long long* totals, uint32_t* order, uint8_t* bins, hipStream_t s);
// Surface voxelization (SVO.cu:37-188): emits composite keys.
template <typename K>
// Hits go to `keys` (raw, for the sort) when `bitmap` is null, else into the
// occupancy bitmap (word bm_off[mesh] + code / 32, bit code % 32, which the
// caller has zeroed); *key_count receives the number of raw hits.
void launch_voxelize(const MeshDev* meshes, const MeshParams* params, const uint64_t* tri_base,
                     uint32_t n_meshes, uint64_t n_tris, int cb, const Tuning& tn, K* keys,
                     uint32_t* bitmap, const uint64_t* bm_off, unsigned long long* key_count,
                     const uint32_t* order, BigItem* big, unsigned long long* big_count,
                     uint64_t big_cap, hipStream_t s);

// ------------------------------------------------------------ k_sort.hip
// LSD radix sort of n (device-resident count) keys on bits [0, bits).
// Returns in *result which of (a, b) holds the sorted keys.  n must not exceed
// cap; `tmp` needs radix_sort_tmp_bytes(cap) bytes.  `blocks` bounds the
// persistent grid of the digit passes.
size_t radix_sort_tmp_bytes(uint64_t cap);
template <typename K>
void launch_radix_sort(K* a, K* b, const unsigned long long* n, uint64_t cap, int bits,
                       uint32_t blocks, void* tmp, K** result, hipStream_t s);

// ------------------------------------------------------------ k_levels.hip
struct LevelArrays {
  void* elem;            // K[cap]: sorted unique keys of this level (S_l)
  uint32_t* elem_group;  // [cap]: group index (at this level) of each element (l >= 2)
  uint32_t* first_child; // [cap]: for groups of the level below: first member index
  uint64_t cap;
};
// Level 1 from the sorted hit keys: unique codes, their groups, and the
// level-2 element array (parents).  Counts go to counts[0] (level 1 elements)
// and counts[1] (level 2 elements).  n_sorted <= cap; `blocks` bounds the
// persistent grid; `tmp` needs level_tmp_bytes(cap) bytes and is shared with
// the later levels of the same pass (launch_level_first prepares it).
template <typename K>
void launch_level_first(const K* sorted, const unsigned long long* n_sorted, uint64_t cap,
                        KeyLayout<K> kl, LevelArrays l1, LevelArrays l2,
                        unsigned long long* counts, LevelSpan* spans1, uint32_t n_levels,
                        uint32_t blocks, void* tmp, hipStream_t s);
// Level 1 from the occupancy bitmap (n_words words; mesh m owns words
// [bm_off[m], bm_off[m+1]), and every bm_off[m] is a multiple of
// BITMAP_TILE_WORDS): the same outputs as launch_level_first followed
// by launch_level_first_finalize, and the bitmap is left all zero (the bytes
// that takes are added to *reset_bytes).  `tmp` needs
// bitmap_level_tmp_bytes(n_words); `lvl_tmp` (level_tmp_bytes of l1.cap) is
// prepared for the later levels.
size_t bitmap_level_tmp_bytes(uint64_t n_words);
// Words per tile of the bitmap's level-1 pass; each mesh's bitmap region
// starts on a tile boundary, so that no tile holds two meshes' words.
constexpr uint64_t BITMAP_TILE_WORDS = 2048;
inline uint64_t bitmap_region_words(uint64_t codes) {
  return ((codes + 31) / 32 + BITMAP_TILE_WORDS - 1) / BITMAP_TILE_WORDS * BITMAP_TILE_WORDS;
}
template <typename K>
void launch_level_first_bitmap(uint32_t* bits, uint64_t n_words, const uint64_t* bm_off,
                               uint32_t n_meshes, KeyLayout<K> kl, LevelArrays l1,
                               LevelArrays l2, unsigned long long* counts, LevelSpan* spans1,
                               uint32_t n_levels, MeshTreeState* state, void* tmp,
                               void* lvl_tmp, unsigned long long* reset_bytes, hipStream_t s);
// Marks meshes whose level-1 set holds a single voxel as depth-1 trees (the
// sort path; launch_level_first_bitmap does it itself).
void launch_level_first_finalize(const LevelSpan* spans, uint32_t n_meshes, uint32_t n_levels,
                                 MeshTreeState* state, hipStream_t s);
// One level l >= 2: roots of meshes whose level holds one element; groups of
// the others feed level l+1.
template <typename K>
void launch_level(int level, KeyLayout<K> kl, LevelArrays cur, LevelArrays next,
                  unsigned long long* counts, LevelSpan* spans, uint32_t n_levels,
                  MeshTreeState* state, uint32_t blocks, void* tmp, hipStream_t s);
size_t level_tmp_bytes(uint64_t cap);

// ------------------------------------------------------------ k_attrs.hip
// Per mesh node layout computed on the host after the level phase.
struct MeshNodeLayout {
  uint64_t node_base;    // first node of the mesh in the batch arrays
  uint64_t n_nodes;
  uint32_t depth;
  uint32_t key_out_bits; // 32 or 64
  uint64_t level_off[24];// node offset of each level within the mesh (level l at [l-1])
  float width[24];       // node width per level
};

struct NodeArrays {
  void* code;            // uint32 or uint64 per node (code_bytes)
  uint8_t* is_leaf;
  float* origin;         // 3 per node
  float* width;
  uint32_t* parent;
  uint32_t* childs;      // 8 per node
  uint32_t* neighbors;   // 27 per node
  int code_bytes;
  // Nonzero: the node and root kernels also set each node's 27 neighbor
  // slots to NONE; zero: a fill kernel does that before them.
  int fill_neighbors;
  // Graph-captured passes: when *skip is nonzero (a size differs from the
  // recorded run) the node, neighbor and record kernels return at once, so
  // that nothing is written past the recorded layout.  Null otherwise.
  const uint32_t* skip;
};

template <typename K>
void launch_emit_level_nodes(int level, KeyLayout<K> kl, const void* elem_l, uint64_t n_elem_l,
                             const void* elem_next, const uint32_t* elem_group_next,
                             const uint32_t* first_child_next, uint64_t n_groups,
                             const LevelSpan* spans, uint32_t n_levels,
                             const MeshTreeState* state, const MeshNodeLayout* layout,
                             const MeshParams* params, NodeArrays out, hipStream_t s);
template <typename K>
void launch_emit_roots(KeyLayout<K> kl, void* const* level_elems, uint32_t n_meshes,
                       const LevelSpan* spans, uint32_t n_levels, const MeshTreeState* state,
                       const MeshNodeLayout* layout, const MeshParams* params, int semantics,
                       NodeArrays out, hipStream_t s);
template <typename K>
void launch_copy_voxels(KeyLayout<K> kl, const void* elem1, uint64_t n, void* out,
                        int out_bytes, hipStream_t s);
void launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s);
// Sets *flag (atomic OR of 1) when the `bytes` bytes (a multiple of 4, both
// 4-byte aligned) at `got` differ from those at `want`.
void launch_check_equal(const void* got, const void* want, uint64_t bytes, uint32_t* flag,
                        hipStream_t s);
template <typename K>
void launch_neighbors_level(int level, KeyLayout<K> kl, const void* elem_next,
                            const uint32_t* elem_group_next, uint64_t n_groups,
                            const LevelSpan* spans, uint32_t n_levels,
                            const MeshTreeState* state, const MeshNodeLayout* layout,
                            NodeArrays nodes, hipStream_t s);
// Vertex / edge / face records: count, scan, write.
struct AttrOut {
  svo_vertex_t* verts;
  svo_edge_t* edges;
  svo_edge_t* faces;
  unsigned long long* totals;  // [3] totals after the write
};
size_t attr_tmp_bytes(uint32_t blocks);
void launch_attrs_count(int semantics, const NodeArrays& nodes, uint64_t n_nodes,
                        const uint64_t* mesh_node_base, uint32_t n_meshes, uint32_t blocks,
                        void* tmp, AttrOut out, hipStream_t s);
void launch_attrs_write(int semantics, const NodeArrays& nodes, uint64_t n_nodes,
                        const uint64_t* mesh_node_base, uint32_t n_meshes, uint32_t blocks,
                        void* tmp, AttrOut out, hipStream_t s);
// Reference-semantics records at identity offsets (8 and 12 per node), valid
// when no record can equal its predecessor; also writes the totals.
void launch_ref_attrs(const NodeArrays& nodes, uint64_t n_nodes, const uint64_t* mesh_node_base,
                      uint32_t n_meshes, AttrOut out, hipStream_t s);
// 168-byte SVONode records for nodes [first, first + n) (32-bit codes).
void launch_pack_aos(const NodeArrays& nodes, uint64_t first, uint64_t n, void* dst,
                     hipStream_t s);
void launch_query_neighbors(const uint32_t* neighbors, uint64_t n_nodes, const uint32_t* node,
                            const uint8_t* slot, uint64_t n, uint32_t* out, hipStream_t s);

}  // namespace svo
