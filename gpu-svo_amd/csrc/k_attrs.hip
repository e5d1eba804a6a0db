// k_attrs.hip -- node records, neighbors and vertex / edge / face records.
//
// Node records (createNode_1 / createNode_2 / createRemainNode,
// SVO.cu:237-339) are written one thread per node, eight threads per parent
// group, after the level phase has fixed every node count:
//   code    = (group parent << 3) | k
//   origin  = gridOrigin + width * (float)decode(code)   per axis
//   isLeaf  = false only for occupied nodes above level 1
//   childs  = child level offset + 8 * (rank of the node's code in its level) + c
//   parent  = the parent group's node in the level above
// The root (the single-code level) is written separately.  Node arrays are
// structure-of-arrays; the reference's 168-byte records are produced on copy.
//
// Neighbors (SVO.cu:522-565).  REFERENCE: the stored table holds UINT_MAX
// except root[13] = root (findNeighbors<false> writes a local copy,
// SVO.cu:535-542).  INTENDED: top-down, nb[s] = childs of the parent-level
// neighbor given by the neighbor LUTs (closed form of MortonLUT.h:531-597).
//
// Vertex / edge records (determineNodeVertex / determineNodeEdge +
// thrust::unique, SVO.cu:586-728).  REFERENCE: 8 corners w*(i&1, i>>1&1,
// i>>2&1) and 12 absolute edges per node, owner = the last stored neighbor of
// the sharing cells (600-603, 669-672), which is the node itself because the
// stored table has no other entries (root[13] is the root); a record equal to
// its predecessor is dropped (uniqueVert / uniqueEdge).  INTENDED: a vertex,
// edge or face is kept by the same-level sharing node with the smallest
// index.  Count -> scan -> write, node-major.
#include <algorithm>

#include "svo_internal.h"
#include "svo_scan.h"

namespace svo {
namespace {

constexpr int BLOCK = 256;

struct LevelPtrs {
  const void* elem[24];
};

template <typename K>
__device__ inline void decode_origin(const MeshParams& P, float w, K code, float o[3]) {
  uint32_t x, y, z;
  demorton3<K>(code, x, y, z);
  o[0] = P.origin[0] + w * (float)x;
  o[1] = P.origin[1] + w * (float)y;
  o[2] = P.origin[2] + w * (float)z;
}

__device__ inline void store_code(const NodeArrays& out, uint64_t t, uint64_t c) {
  if (out.code_bytes == 8) ((uint64_t*)out.code)[t] = c;
  else ((uint32_t*)out.code)[t] = (uint32_t)c;
}

__device__ inline void store_childs(uint32_t* ch, uint32_t base, bool valid) {
  uint4 a, b;
  if (valid) {
    a = make_uint4(base, base + 1, base + 2, base + 3);
    b = make_uint4(base + 4, base + 5, base + 6, base + 7);
  } else {
    a = make_uint4(NONE, NONE, NONE, NONE);
    b = a;
  }
  reinterpret_cast<uint4*>(ch)[0] = a;
  reinterpret_cast<uint4*>(ch)[1] = b;
}

// Node (group g at `level`, child key k): shared by the node and neighbor kernels.
template <typename K>
struct GroupNode {
  bool live;
  uint32_t m;
  K P;            // parent key (element g of the next level)
  uint64_t local; // node index within its mesh
  uint32_t parent;
};

template <typename K>
__device__ inline GroupNode<K> group_node(int level, KeyLayout<K> kl, const K* elem_next,
                                          const uint32_t* group_next, uint64_t n_groups,
                                          const LevelSpan* spans, uint32_t n_levels,
                                          const MeshTreeState* state,
                                          const MeshNodeLayout* layout, uint64_t g, uint32_t k) {
  GroupNode<K> r;
  r.live = false;
  r.m = 0;
  r.P = 0;
  r.local = 0;
  r.parent = NONE;
  if (g >= n_groups) return r;
  r.P = elem_next[g];
  r.m = kl.mesh(r.P);
  const uint32_t depth = state[r.m].depth;
  if (depth <= (uint32_t)level) return r;  // depth-1 tree: its level-1 group is not a node group
  const MeshNodeLayout& L = layout[r.m];
  if (L.n_nodes == 0) return r;  // mesh rejected after the level phase
  r.live = true;
  const LevelSpan sp = spans[(uint64_t)r.m * n_levels + (level - 1)];
  r.local = L.level_off[level - 1] + 8ull * (g - sp.group_first) + k;
  if (depth == (uint32_t)level + 1) {
    r.parent = (uint32_t)L.level_off[level];  // the root
  } else {
    const LevelSpan spn = spans[(uint64_t)r.m * n_levels + level];
    r.parent = (uint32_t)(L.level_off[level] + 8ull * (group_next[g] - spn.group_first) +
                          (uint32_t)(r.P & 7));
  }
  return r;
}

// Everything a node record needs, gathered in two rounds of loads that do not
// depend on each other within a round: the group arrays, then what is keyed
// by the mesh and by the first child.
template <typename K>
struct NodeIn {
  K P, e;
  uint32_t fc, gnext, depth, spn_first;
  uint64_t g, nn, nbase, off_l, off_p, off_c;
  LevelSpan sp;
  float w, ox, oy, oz;
  bool in;
};

constexpr int NODES_PT = 4;  // nodes per thread, their loads in flight together

// One thread per node, eight per parent group, NODES_PT nodes per thread
// BLOCK apart (so stores stay contiguous across a wave).  All loads of the
// thread's nodes are issued before the first is used.
template <typename K>
__global__ __launch_bounds__(BLOCK) void k_emit_nodes(
    int level, KeyLayout<K> kl, const K* elem_l, uint64_t n_elem_l, const K* elem_next,
    const uint32_t* group_next, const uint32_t* first_child_next, uint64_t n_groups,
    const LevelSpan* spans, uint32_t n_levels, const MeshTreeState* state,
    const MeshNodeLayout* layout, const MeshParams* params, NodeArrays out) {
  if (out.skip && *out.skip) return;  // graph replay over changed sizes
  const uint64_t base = (uint64_t)blockIdx.x * BLOCK * NODES_PT + threadIdx.x;
  const uint32_t k = (uint32_t)(threadIdx.x & 7);
  NodeIn<K> a[NODES_PT];
#pragma unroll
  for (int u = 0; u < NODES_PT; ++u) {  // first loads: the parent group
    const uint64_t g0 = (base + (uint64_t)u * BLOCK) >> 3;
    a[u].in = g0 < n_groups;
    a[u].g = a[u].in ? g0 : 0;
    a[u].P = elem_next[a[u].g];
    a[u].fc = first_child_next[a[u].g];
    a[u].gnext = group_next[a[u].g];
  }
  auto mesh_fields = [&](NodeIn<K>& q, uint32_t m) {
    q.depth = state[m].depth;
    const MeshNodeLayout* L = layout + m;
    q.nn = L->n_nodes;
    q.nbase = L->node_base;
    q.off_l = L->level_off[level - 1];
    q.off_p = L->level_off[level];
    q.off_c = level >= 2 ? L->level_off[level - 2] : 0;
    q.w = L->width[level - 1];
    q.sp = spans[(uint64_t)m * n_levels + (level - 1)];
    q.spn_first = spans[(uint64_t)m * n_levels + level].group_first;
    q.ox = params[m].origin[0];
    q.oy = params[m].origin[1];
    q.oz = params[m].origin[2];
  };
#pragma unroll
  for (int u = 0; u < NODES_PT; ++u) {  // second loads: mesh records, child key
    const uint32_t m = kl.mesh(a[u].P);
    // a wave's nodes nearly always share one mesh: its per-mesh records are
    // then read once for the wave (scalar loads) instead of once per lane
    const uint32_t m0 = __builtin_amdgcn_readfirstlane(m);
    if (__all(m == m0)) mesh_fields(a[u], m0);
    else mesh_fields(a[u], m);
    const uint64_t j = (uint64_t)a[u].fc + k;
    a[u].e = j < n_elem_l ? elem_l[j] : (K)0;
  }
  if (out.fill_neighbors) {
    // the nodes' 27 neighbor slots, all NONE (the roots' slots and the
    // intended semantics' neighbors are written by later kernels), issued
    // while the child keys are still loading: they need the mesh records
    // only.  When the wave's 64 nodes are live and consecutive, their 1728
    // slots are one contiguous run, written by the wave with 16-byte stores.
#pragma unroll
    for (int u = 0; u < NODES_PT; ++u) {
      const NodeIn<K>& q = a[u];
      const bool live = q.in && q.depth > (uint32_t)level && q.nn != 0;
      const uint64_t tn = q.nbase + q.off_l + 8ull * (q.g - q.sp.group_first) + k;
      const uint64_t tb = tn - __lane_id();
      const uint64_t tb0 =
          ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(tb >> 32)) << 32) |
          (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)tb);
      if (__all(live && tb == tb0)) {
        // 1728 slots from a 4-byte boundary: up to 3 single slots to reach
        // a 16-byte boundary, 16-byte stores, then the remaining slots
        uint32_t* nb = out.neighbors + 27 * tb0;
        const uint32_t head = (uint32_t)((4 - (((uintptr_t)nb >> 2) & 3)) & 3);
        const uint32_t lane = __lane_id();
        if (lane < head) nb[lane] = NONE;
        uint4* body = (uint4*)(nb + head);
        const uint32_t n4 = (1728 - head) >> 2;
        const uint4 none4 = make_uint4(NONE, NONE, NONE, NONE);
#pragma unroll
        for (uint32_t i = 0; i < 7; ++i)
          if (lane + 64 * i < n4) body[lane + 64 * i] = none4;
        const uint32_t tail = (1728 - head) & 3;
        if (lane < tail) nb[head + 4 * n4 + lane] = NONE;
      } else if (live) {
        uint32_t* nb = out.neighbors + 27 * tn;
#pragma unroll
        for (int i = 0; i < 27; ++i) nb[i] = NONE;
      }
    }
  }
#pragma unroll
  for (int u = 0; u < NODES_PT; ++u) {
    const NodeIn<K>& q = a[u];
    const uint64_t j = (uint64_t)q.fc + k;
    // a depth-1 tree's level-1 group is no node group; a mesh rejected after
    // the level phase has no nodes
    const bool live = q.in && q.depth > (uint32_t)level && q.nn != 0;
    uint32_t bit =
        (live && j < n_elem_l && kl.parent(q.e) == q.P) ? 1u << (uint32_t)(q.e & 7) : 0u;
    bit |= __shfl_xor(bit, 1);
    bit |= __shfl_xor(bit, 2);
    bit |= __shfl_xor(bit, 4);
    if (!live) continue;
    const uint32_t mask = bit;
    const bool real = (mask >> k) & 1u;
    const uint64_t local = q.off_l + 8ull * (q.g - q.sp.group_first) + k;
    const uint64_t t = q.nbase + local;
    const uint32_t parent = q.depth == (uint32_t)level + 1
                                ? (uint32_t)q.off_p  // the root
                                : (uint32_t)(q.off_p + 8ull * (q.gnext - q.spn_first) +
                                             (uint32_t)(q.P & 7));
    const K code = (kl.code(q.P) << 3) | (K)k;
    uint32_t x, y, z;
    demorton3<K>(code, x, y, z);
    store_code(out, t, (uint64_t)code);
    out.is_leaf[t] = (level == 1 || !real) ? 1 : 0;
    // gridOrigin + width * Vector3f(x, y, z), per axis (SVO.cu:262)
    out.origin[3 * t + 0] = q.ox + q.w * (float)x;
    out.origin[3 * t + 1] = q.oy + q.w * (float)y;
    out.origin[3 * t + 2] = q.oz + q.w * (float)z;
    out.width[t] = q.w;
    out.parent[t] = parent;
    uint32_t cbase = 0;
    const bool has_childs = real && level >= 2;
    if (has_childs) {
      const uint64_t i = (uint64_t)q.fc + __popc(mask & ((1u << k) - 1u));
      cbase = (uint32_t)(q.off_c + 8ull * (i - q.sp.elem_first));
    }
    store_childs(out.childs + 8 * t, cbase, has_childs);
  }
}

template <typename K>
__global__ void k_emit_roots(KeyLayout<K> kl, LevelPtrs lp, uint32_t n_meshes,
                             const LevelSpan* spans, uint32_t n_levels,
                             const MeshTreeState* state, const MeshNodeLayout* layout,
                             const MeshParams* params, NodeArrays out) {
  if (out.skip && *out.skip) return;  // graph replay over changed sizes
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n_meshes) return;
  const uint32_t d = state[m].depth;
  const MeshNodeLayout& L = layout[m];
  if (d == 0 || L.n_nodes == 0) return;
  const uint64_t local = L.level_off[d - 1];
  const uint64_t t = L.node_base + local;
  const K key = ((const K*)lp.elem[d - 1])[state[m].root_elem];
  const K code = kl.code(key);
  const float w = L.width[d - 1];
  float o[3];
  decode_origin<K>(params[m], w, code, o);
  store_code(out, t, (uint64_t)code);
  out.is_leaf[t] = d == 1 ? 1 : 0;
  out.origin[3 * t + 0] = o[0];
  out.origin[3 * t + 1] = o[1];
  out.origin[3 * t + 2] = o[2];
  out.width[t] = w;
  out.parent[t] = NONE;
  store_childs(out.childs + 8 * t, d >= 2 ? (uint32_t)L.level_off[d - 2] : 0u, d >= 2);
  // findNeighbors<true> (SVO.cu:528-529): root.neighbors[13] = root index,
  // every other slot NONE (written here when no fill kernel ran)
  if (out.fill_neighbors) {
    for (int i = 0; i < 27; ++i)
      out.neighbors[27 * t + i] = (i == 13 && d >= 2) ? (uint32_t)local : NONE;
  } else if (d >= 2) {
    out.neighbors[27 * t + 13] = (uint32_t)local;
  }
  (void)spans;
  (void)n_levels;
}

template <typename K>
__global__ void k_copy_voxels(KeyLayout<K> kl, const K* elem1, uint64_t n, void* out,
                              int out_bytes) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const K c = kl.code(elem1[i]);
  if (out_bytes == 8) ((uint64_t*)out)[i] = (uint64_t)c;
  else ((uint32_t*)out)[i] = (uint32_t)c;
}

// Grid-stride (n may exceed what one launch can index: 108 bytes of
// neighbors per node are 2^32 words at 160M nodes).  p is 16-byte aligned.
__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * 4;
  for (uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4; i < n; i += stride) {
    if (i + 3 < n) {
      store_stream(reinterpret_cast<uint4*>(p + i), make_uint4(v, v, v, v));
    } else {
      for (uint64_t j = i; j < n; ++j) p[j] = v;
    }
  }
}

// Neighbor LUTs (MortonLUT.h:531-597) in closed form: slot s = (dx+1) +
// 3(dy+1) + 9(dz+1); child k's neighbor in direction d is child
// ((k+d) mod 2 per axis) of the parent-level neighbor floor((k+d)/2).
__device__ inline void nb_lut(uint32_t k, uint32_t s, uint32_t& pslot, uint32_t& child) {
  const int d[3] = {(int)(s % 3) - 1, (int)((s / 3) % 3) - 1, (int)(s / 9) - 1};
  const int mul[3] = {1, 3, 9};
  int ps = 0, ch = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int c = (int)((k >> a) & 1u) + d[a];
    const int p = c < 0 ? -1 : (c >> 1);
    ps += (p + 1) * mul[a];
    ch |= (c & 1) << a;
  }
  pslot = (uint32_t)ps;
  child = (uint32_t)ch;
}

template <typename K>
__global__ __launch_bounds__(BLOCK) void k_neighbors_level(
    int level, KeyLayout<K> kl, const K* elem_next, const uint32_t* group_next,
    uint64_t n_groups, const LevelSpan* spans, uint32_t n_levels, const MeshTreeState* state,
    const MeshNodeLayout* layout, NodeArrays nodes) {
  if (nodes.skip && *nodes.skip) return;  // graph replay over changed sizes
  const uint64_t gid = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  const uint64_t g = gid >> 3;
  const uint32_t k = (uint32_t)(gid & 7);
  const GroupNode<K> gn = group_node(level, kl, elem_next, group_next, n_groups, spans, n_levels,
                                     state, layout, g, k);
  if (!gn.live) return;
  const uint64_t base = layout[gn.m].node_base;
  const uint32_t* pn_row = nodes.neighbors + 27 * (base + gn.parent);
  uint32_t* row = nodes.neighbors + 27 * (base + gn.local);
  for (uint32_t s = 0; s < 27; ++s) {
    uint32_t ps, ch;
    nb_lut(k, s, ps, ch);
    const uint32_t pn = pn_row[ps];
    row[s] = pn != NONE ? nodes.childs[8 * (base + pn) + ch] : NONE;
  }
}

// ---------------------------------------------------------------- vertex / edge / face
// Edge endpoints in units of the node width, in the order of SVO.cu:638-663.
__constant__ uint8_t EDGE_OFS[12][2][3] = {
    {{0, 0, 0}, {0, 1, 0}}, {{0, 1, 0}, {1, 1, 0}}, {{1, 1, 0}, {1, 0, 0}}, {{1, 0, 0}, {0, 0, 0}},
    {{0, 0, 1}, {0, 1, 1}}, {{0, 1, 1}, {1, 1, 1}}, {{1, 1, 1}, {1, 0, 1}}, {{1, 0, 1}, {0, 0, 1}},
    {{0, 0, 0}, {0, 0, 1}}, {{0, 1, 0}, {0, 1, 1}}, {{1, 1, 0}, {1, 1, 1}}, {{1, 0, 0}, {1, 0, 1}}};
// Faces: -x, +x, -y, +y, -z, +z as (lo corner, hi corner) and the cell across.
__constant__ uint8_t FACE_OFS[6][2][3] = {{{0, 0, 0}, {0, 1, 1}}, {{1, 0, 0}, {1, 1, 1}},
                                          {{0, 0, 0}, {1, 0, 1}}, {{0, 1, 0}, {1, 1, 1}},
                                          {{0, 0, 0}, {1, 1, 0}}, {{0, 0, 1}, {1, 1, 1}}};
__constant__ uint8_t FACE_SLOT[6] = {12, 14, 10, 16, 4, 22};

// origin + Vector3f(0|w, 0|w, 0|w), with the bare origin for the zero offset
// (SVO.cu:640-662).
__device__ inline void offset_point(const float* o, float w, const uint8_t* ofs, float* p) {
  const bool bare = !ofs[0] && !ofs[1] && !ofs[2];
#pragma unroll
  for (int a = 0; a < 3; ++a) p[a] = bare ? o[a] : o[a] + (ofs[a] ? w : 0.0f);
}

__device__ inline bool eq3(const float* a, const float* b) {
  return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

struct NodeGeom {
  float o[3];
  float w;
};
__device__ inline NodeGeom load_geom(const NodeArrays& n, uint64_t t) {
  NodeGeom g;
  g.o[0] = n.origin[3 * t];
  g.o[1] = n.origin[3 * t + 1];
  g.o[2] = n.origin[3 * t + 2];
  g.w = n.width[t];
  return g;
}

__device__ inline void ref_corner(float w, int j, float* p) {
  p[0] = w * (float)(j & 1);
  p[1] = w * (float)((j >> 1) & 1);
  p[2] = w * (float)((j >> 2) & 1);
}
__device__ inline void ref_edge(const NodeGeom& g, int i, float* a, float* b) {
  offset_point(g.o, g.w, EDGE_OFS[i][0], a);
  offset_point(g.o, g.w, EDGE_OFS[i][1], b);
}
__device__ inline bool edge_eq(const float* a0, const float* b0, const float* a1, const float* b1) {
  return (eq3(a0, a1) && eq3(b0, b1)) || (eq3(a0, b1) && eq3(b0, a1));
}

// Cells sharing corner j: d in {b-1, b} per axis, x fastest (SVO.cu:567-584).
__device__ inline uint32_t vert_slot(int j, int q) {
  const int dx = ((j >> 0) & 1) - 1 + ((q >> 0) & 1);
  const int dy = ((j >> 1) & 1) - 1 + ((q >> 1) & 1);
  const int dz = ((j >> 2) & 1) - 1 + ((q >> 2) & 1);
  return (uint32_t)((dx + 1) + 3 * (dy + 1) + 9 * (dz + 1));
}
// Cells sharing edge i: the two axes the edge does not run along take
// {o-1, o}; the running axis takes 0 (SVO.cu:610-627).
__device__ inline uint32_t edge_slot(int i, int q) {
  const uint8_t* a = EDGE_OFS[i][0];
  const uint8_t* b = EDGE_OFS[i][1];
  int d[3];
  int bitpos = 0;
  for (int ax = 0; ax < 3; ++ax) {
    if (a[ax] != b[ax]) {
      d[ax] = 0;
    } else {
      d[ax] = (int)a[ax] - 1 + ((q >> bitpos) & 1);
      ++bitpos;
    }
  }
  return (uint32_t)((d[0] + 1) + 3 * (d[1] + 1) + 9 * (d[2] + 1));
}

__device__ inline uint32_t find_mesh64(const uint64_t* base, uint32_t n_meshes, uint64_t t) {
  uint32_t lo = 0, hi = n_meshes;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (base[mid] <= t) lo = mid; else hi = mid;
  }
  return lo;
}

struct AttrCounts {
  uint32_t v, e, f;
};

// Counts (and optionally writes) the records of node t.
template <bool WRITE>
__device__ inline AttrCounts node_attrs(int semantics, const NodeArrays& n, uint64_t t,
                                        uint64_t local, uint64_t vpos, uint64_t epos,
                                        uint64_t fpos, AttrOut out) {
  AttrCounts c = {0, 0, 0};
  const NodeGeom g = load_geom(n, t);
  const uint32_t self = (uint32_t)local;
  if (semantics == SVO_SEM_REFERENCE) {
    float prev[3];
    if (local > 0) ref_corner(n.width[t - 1], 7, prev);
    for (int j = 0; j < 8; ++j) {
      float p[3];
      ref_corner(g.w, j, p);
      const bool keep = (j == 0 && local == 0) || !eq3(p, prev);
      if (keep) {
        if (WRITE) {
          svo_vertex_t r;
          r.p[0] = p[0];
          r.p[1] = p[1];
          r.p[2] = p[2];
          r.owner = self;
          out.verts[vpos + c.v] = r;
        }
        ++c.v;
      }
      prev[0] = p[0];
      prev[1] = p[1];
      prev[2] = p[2];
    }
    float pa[3], pb[3];
    if (local > 0) {
      const NodeGeom gp = load_geom(n, t - 1);
      ref_edge(gp, 11, pa, pb);
    }
    for (int i = 0; i < 12; ++i) {
      float a[3], b[3];
      ref_edge(g, i, a, b);
      const bool keep = (i == 0 && local == 0) || !edge_eq(a, b, pa, pb);
      if (keep) {
        if (WRITE) {
          svo_edge_t r;
          for (int q = 0; q < 3; ++q) {
            r.a[q] = a[q];
            r.b[q] = b[q];
          }
          r.owner = self;
          out.edges[epos + c.e] = r;
        }
        ++c.e;
      }
      for (int q = 0; q < 3; ++q) {
        pa[q] = a[q];
        pb[q] = b[q];
      }
    }
    return c;
  }
  const uint32_t* nb = n.neighbors + 27 * t;
  for (int j = 0; j < 8; ++j) {
    uint32_t owner = self;
    for (int q = 0; q < 8; ++q) {
      const uint32_t v = nb[vert_slot(j, q)];
      owner = v < owner ? v : owner;
    }
    if (owner != self) continue;
    if (WRITE) {
      const uint8_t ofs[3] = {(uint8_t)(j & 1), (uint8_t)((j >> 1) & 1), (uint8_t)((j >> 2) & 1)};
      svo_vertex_t r;
      offset_point(g.o, g.w, ofs, r.p);
      r.owner = self;
      out.verts[vpos + c.v] = r;
    }
    ++c.v;
  }
  for (int i = 0; i < 12; ++i) {
    uint32_t owner = self;
    for (int q = 0; q < 4; ++q) {
      const uint32_t v = nb[edge_slot(i, q)];
      owner = v < owner ? v : owner;
    }
    if (owner != self) continue;
    if (WRITE) {
      svo_edge_t r;
      ref_edge(g, i, r.a, r.b);
      r.owner = self;
      out.edges[epos + c.e] = r;
    }
    ++c.e;
  }
  for (int f = 0; f < 6; ++f) {
    const uint32_t v = nb[FACE_SLOT[f]];
    if (v < self) continue;
    if (WRITE) {
      svo_edge_t r;
      offset_point(g.o, g.w, FACE_OFS[f][0], r.a);
      offset_point(g.o, g.w, FACE_OFS[f][1], r.b);
      r.owner = self;
      out.faces[fpos + c.f] = r;
    }
    ++c.f;
  }
  return c;
}

__device__ inline void block_range(uint64_t n, uint32_t G, uint32_t b, uint64_t& beg,
                                   uint64_t& end) {
  const uint64_t chunk = (n + G - 1) / G;
  beg = (uint64_t)b * chunk;
  if (beg > n) beg = n;
  end = beg + chunk;
  if (end > n) end = n;
}

__global__ __launch_bounds__(BLOCK) void k_attr_count(int semantics, NodeArrays nodes,
                                                      uint64_t n_nodes, const uint64_t* mbase,
                                                      uint32_t n_meshes, uint32_t G,
                                                      uint64_t* partial) {
  if (nodes.skip && *nodes.skip) return;  // graph replay over changed sizes
  __shared__ uint64_t tmp[4];
  uint64_t beg, end;
  block_range(n_nodes, G, blockIdx.x, beg, end);
  uint64_t sv = 0, se = 0, sf = 0;
  AttrOut none = {nullptr, nullptr, nullptr, nullptr};
  uint32_t m = beg < end ? find_mesh64(mbase, n_meshes, beg) : 0;
  for (uint64_t t = beg + threadIdx.x; t < end; t += BLOCK) {
    while (m + 1 < n_meshes && mbase[m + 1] <= t) ++m;
    const AttrCounts c = node_attrs<false>(semantics, nodes, t, t - mbase[m], 0, 0, 0, none);
    sv += c.v;
    se += c.e;
    sf += c.f;
  }
  uint64_t tv, te, tf;
  block_excl_scan256(sv, tmp, tv);
  block_excl_scan256(se, tmp, te);
  block_excl_scan256(sf, tmp, tf);
  if (threadIdx.x == 0) {
    partial[3 * blockIdx.x] = tv;
    partial[3 * blockIdx.x + 1] = te;
    partial[3 * blockIdx.x + 2] = tf;
  }
}

__global__ __launch_bounds__(BLOCK) void k_triple_scan(uint64_t* partial, uint32_t G,
                                                       unsigned long long* totals) {
  __shared__ uint64_t tmp[4];
  const uint32_t per = (G + BLOCK - 1) / BLOCK;
  const uint32_t b0 = threadIdx.x * per;
  for (int c = 0; c < 3; ++c) {
    uint64_t s = 0;
    for (uint32_t k = 0; k < per; ++k)
      if (b0 + k < G) s += partial[3 * (b0 + k) + c];
    uint64_t tot;
    uint64_t run = block_excl_scan256(s, tmp, tot);
    for (uint32_t k = 0; k < per; ++k)
      if (b0 + k < G) {
        const uint64_t v = partial[3 * (b0 + k) + c];
        partial[3 * (b0 + k) + c] = run;
        run += v;
      }
    if (threadIdx.x == 0) totals[c] = tot;
  }
}

__global__ __launch_bounds__(BLOCK) void k_attr_write(int semantics, NodeArrays nodes,
                                                      uint64_t n_nodes, const uint64_t* mbase,
                                                      uint32_t n_meshes, uint32_t G,
                                                      const uint64_t* partial, AttrOut out,
                                                      uint64_t* mesh_first) {
  if (nodes.skip && *nodes.skip) return;  // graph replay over changed sizes
  __shared__ uint64_t tmp[4];
  uint64_t beg, end;
  block_range(n_nodes, G, blockIdx.x, beg, end);
  uint64_t bv = partial[3 * blockIdx.x], be = partial[3 * blockIdx.x + 1],
           bf = partial[3 * blockIdx.x + 2];
  AttrOut none = {nullptr, nullptr, nullptr, nullptr};
  uint32_t m = beg < end ? find_mesh64(mbase, n_meshes, beg) : 0;
  for (uint64_t tb = beg; tb < end; tb += BLOCK) {
    const uint64_t t = tb + threadIdx.x;
    AttrCounts c = {0, 0, 0};
    uint64_t local = 0;
    if (t < end) {
      while (m + 1 < n_meshes && mbase[m + 1] <= t) ++m;
      local = t - mbase[m];
      c = node_attrs<false>(semantics, nodes, t, local, 0, 0, 0, none);
    }
    uint64_t tv, te, tf;
    const uint64_t ov = block_excl_scan256((uint64_t)c.v, tmp, tv) + bv;
    const uint64_t oe = block_excl_scan256((uint64_t)c.e, tmp, te) + be;
    const uint64_t of = block_excl_scan256((uint64_t)c.f, tmp, tf) + bf;
    if (t < end) {
      node_attrs<true>(semantics, nodes, t, local, ov, oe, of, out);
      if (local == 0) {
        mesh_first[3 * m] = ov;
        mesh_first[3 * m + 1] = oe;
        mesh_first[3 * m + 2] = of;
      }
    }
    bv += tv;
    be += te;
    bf += tf;
  }
}

// Reference records when no record can equal its predecessor (the host has
// checked that every width is positive, finite and not absorbed by the
// coordinates, so all corners of a node are distinct and thrust::unique keeps
// everything): node t's records sit at the identity offsets 8t and 12t.
// A block handles 64 consecutive nodes: four lanes per node build two
// vertices and three edges each into LDS, then the block streams its
// contiguous 8 KiB of vertex records and 21 KiB of edge records out as
// aligned 16-byte stores (64 * 336 bytes and 64 * 128 bytes are multiples
// of 16).  The mesh of the first node is found once per block.
constexpr int REF_NODES = 64;

__device__ inline void stream_out(const uint32_t* src, uint32_t n_dw, uint32_t* dst) {
  const uint32_t nq = n_dw >> 2;
  for (uint32_t q = threadIdx.x; q < nq; q += BLOCK)
    store_stream(reinterpret_cast<uint4*>(dst) + q, reinterpret_cast<const uint4*>(src)[q]);
  for (uint32_t d = (nq << 2) + threadIdx.x; d < n_dw; d += BLOCK) dst[d] = src[d];
}

__global__ __launch_bounds__(BLOCK) void k_ref_records(const float* origin, const float* width,
                                                       uint64_t n_nodes, const uint64_t* mbase,
                                                       uint32_t n_meshes, uint32_t* verts,
                                                       uint32_t* edges) {
  __shared__ __align__(16) uint32_t sv[REF_NODES * 32];
  __shared__ __align__(16) uint32_t se[REF_NODES * 84];
  __shared__ uint32_t s_mesh;
  const uint64_t first = (uint64_t)blockIdx.x * REF_NODES;
  if (first >= n_nodes) return;
  const uint32_t nn = (uint32_t)((n_nodes - first) < REF_NODES ? (n_nodes - first) : REF_NODES);
  if (threadIdx.x == 0) s_mesh = find_mesh64(mbase, n_meshes, first);
  __syncthreads();
  const uint32_t ln = threadIdx.x >> 2, part = threadIdx.x & 3;
  if (ln < nn) {
    const uint64_t t = first + ln;
    uint32_t m = s_mesh;
    while (m + 1 < n_meshes && mbase[m + 1] <= t) ++m;
    const uint32_t self = (uint32_t)(t - mbase[m]);
    NodeGeom g;
    g.o[0] = origin[3 * t];
    g.o[1] = origin[3 * t + 1];
    g.o[2] = origin[3 * t + 2];
    g.w = width[t];
    for (int j = 2 * part; j < 2 * part + 2; ++j) {
      float p[3];
      ref_corner(g.w, j, p);
      uint32_t* d = sv + 32 * ln + 4 * j;
      d[0] = __float_as_uint(p[0]);
      d[1] = __float_as_uint(p[1]);
      d[2] = __float_as_uint(p[2]);
      d[3] = self;
    }
    for (int i = 3 * part; i < 3 * part + 3; ++i) {
      float a[3], b[3];
      ref_edge(g, i, a, b);
      uint32_t* d = se + 84 * ln + 7 * i;
      d[0] = __float_as_uint(a[0]);
      d[1] = __float_as_uint(a[1]);
      d[2] = __float_as_uint(a[2]);
      d[3] = __float_as_uint(b[0]);
      d[4] = __float_as_uint(b[1]);
      d[5] = __float_as_uint(b[2]);
      d[6] = self;
    }
  }
  __syncthreads();
  stream_out(sv, 32 * nn, verts + 32 * first);
  stream_out(se, 84 * nn, edges + 84 * first);
}

__global__ void k_ref_totals(uint64_t n_nodes, const uint64_t* mbase, uint32_t n_meshes,
                             unsigned long long* totals) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m == 0) {
    totals[0] = 8 * n_nodes;
    totals[1] = 12 * n_nodes;
    totals[2] = 0;
  }
  if (m < n_meshes) {
    totals[4 + 3 * m] = 8 * mbase[m];
    totals[4 + 3 * m + 1] = 12 * mbase[m];
    totals[4 + 3 * m + 2] = 0;
  }
}

// 168-byte SVONode records (SVO.h:20-35), one thread per 32-bit word.
__global__ void k_pack_aos(NodeArrays n, uint64_t first, uint64_t count, uint32_t* dst) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= count * 42) return;
  const uint64_t r = gid / 42;
  const uint32_t wd = (uint32_t)(gid - r * 42);
  const uint64_t t = first + r;
  uint32_t v;
  if (wd == 0) v = n.code_bytes == 8 ? (uint32_t)((const uint64_t*)n.code)[t]
                                     : ((const uint32_t*)n.code)[t];
  else if (wd == 1) v = n.is_leaf[t];  // bool + three zero padding bytes
  else if (wd < 5) v = __float_as_uint(n.origin[3 * t + (wd - 2)]);
  else if (wd == 5) v = __float_as_uint(n.width[t]);
  else if (wd == 6) v = n.parent[t];
  else if (wd < 15) v = n.childs[8 * t + (wd - 7)];
  else v = n.neighbors[27 * t + (wd - 15)];
  dst[gid] = v;
}

__global__ void k_query(const uint32_t* nb, uint64_t n_nodes, const uint32_t* node,
                        const uint8_t* slot, uint64_t n, uint32_t* out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t t = node[i];
  const uint32_t s = slot[i];
  out[i] = (t < n_nodes && s < 27) ? nb[27 * (uint64_t)t + s] : NONE;
}

inline uint32_t grid_for(uint64_t threads) { return (uint32_t)((threads + BLOCK - 1) / BLOCK); }

}  // namespace

template <typename K>
void launch_emit_level_nodes(int level, KeyLayout<K> kl, const void* elem_l, uint64_t n_elem_l,
                             const void* elem_next, const uint32_t* elem_group_next,
                             const uint32_t* first_child_next, uint64_t n_groups,
                             const LevelSpan* spans, uint32_t n_levels,
                             const MeshTreeState* state, const MeshNodeLayout* layout,
                             const MeshParams* params, NodeArrays out, hipStream_t s) {
  if (n_groups == 0) return;
  hipLaunchKernelGGL(k_emit_nodes<K>, dim3(grid_for((8 * n_groups + NODES_PT - 1) / NODES_PT)),
                     dim3(BLOCK), 0, s, level, kl,
                     (const K*)elem_l, n_elem_l, (const K*)elem_next, elem_group_next,
                     first_child_next, n_groups, spans, n_levels, state, layout, params, out);
}

template <typename K>
void launch_emit_roots(KeyLayout<K> kl, void* const* level_elems, uint32_t n_meshes,
                       const LevelSpan* spans, uint32_t n_levels, const MeshTreeState* state,
                       const MeshNodeLayout* layout, const MeshParams* params, int semantics,
                       NodeArrays out, hipStream_t s) {
  (void)semantics;
  LevelPtrs lp;
  for (uint32_t l = 0; l < 24; ++l) lp.elem[l] = l < n_levels ? level_elems[l] : nullptr;
  hipLaunchKernelGGL(k_emit_roots<K>, dim3((n_meshes + 63) / 64), dim3(64), 0, s, kl, lp,
                     n_meshes, spans, n_levels, state, layout, params, out);
}

template <typename K>
void launch_copy_voxels(KeyLayout<K> kl, const void* elem1, uint64_t n, void* out,
                        int out_bytes, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_copy_voxels<K>, dim3(grid_for(n)), dim3(BLOCK), 0, s, kl,
                     (const K*)elem1, n, out, out_bytes);
}

// Sets *flag when any of the n words differs from its expected value.
__global__ void k_check_equal(const uint32_t* got, const uint32_t* want, uint64_t n,
                              uint32_t* flag) {
  uint32_t diff = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    diff |= got[i] ^ want[i];
  if (__any(diff != 0) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

void launch_check_equal(const void* got, const void* want, uint64_t bytes, uint32_t* flag,
                        hipStream_t s) {
  const uint64_t n = bytes / 4;
  if (n == 0) return;
  const uint64_t blocks = std::min<uint64_t>((n + BLOCK - 1) / BLOCK, 256);
  hipLaunchKernelGGL(k_check_equal, dim3((uint32_t)blocks), dim3(BLOCK), 0, s,
                     (const uint32_t*)got, (const uint32_t*)want, n, flag);
}

void launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  const uint64_t blocks = std::min<uint64_t>(grid_for((n + 3) / 4), 1u << 20);
  hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(BLOCK), 0, s, p, value, n);
}

template <typename K>
void launch_neighbors_level(int level, KeyLayout<K> kl, const void* elem_next,
                            const uint32_t* elem_group_next, uint64_t n_groups,
                            const LevelSpan* spans, uint32_t n_levels,
                            const MeshTreeState* state, const MeshNodeLayout* layout,
                            NodeArrays nodes, hipStream_t s) {
  if (n_groups == 0) return;
  hipLaunchKernelGGL(k_neighbors_level<K>, dim3(grid_for(8 * n_groups)), dim3(BLOCK), 0, s,
                     level, kl, (const K*)elem_next, elem_group_next, n_groups, spans, n_levels,
                     state, layout, nodes);
}

size_t attr_tmp_bytes(uint32_t blocks) { return (size_t)3 * blocks * sizeof(uint64_t); }

void launch_attrs_count(int semantics, const NodeArrays& nodes, uint64_t n_nodes,
                        const uint64_t* mesh_node_base, uint32_t n_meshes, uint32_t blocks,
                        void* tmp, AttrOut out, hipStream_t s) {
  uint64_t* partial = (uint64_t*)tmp;
  hipLaunchKernelGGL(k_attr_count, dim3(blocks), dim3(BLOCK), 0, s, semantics, nodes, n_nodes,
                     mesh_node_base, n_meshes, blocks, partial);
  hipLaunchKernelGGL(k_triple_scan, dim3(1), dim3(BLOCK), 0, s, partial, blocks, out.totals);
}

void launch_attrs_write(int semantics, const NodeArrays& nodes, uint64_t n_nodes,
                        const uint64_t* mesh_node_base, uint32_t n_meshes, uint32_t blocks,
                        void* tmp, AttrOut out, hipStream_t s) {
  uint64_t* partial = (uint64_t*)tmp;
  uint64_t* mesh_first = (uint64_t*)out.totals + 4;  // [3 * n_meshes] after the totals
  hipLaunchKernelGGL(k_attr_write, dim3(blocks), dim3(BLOCK), 0, s, semantics, nodes, n_nodes,
                     mesh_node_base, n_meshes, blocks, (const uint64_t*)partial, out, mesh_first);
}

void launch_ref_attrs(const NodeArrays& nodes, uint64_t n_nodes, const uint64_t* mesh_node_base,
                      uint32_t n_meshes, AttrOut out, hipStream_t s) {
  hipLaunchKernelGGL(k_ref_totals, dim3((n_meshes + 255) / 256), dim3(BLOCK), 0, s, n_nodes,
                     mesh_node_base, n_meshes, out.totals);
  if (n_nodes == 0) return;
  hipLaunchKernelGGL(k_ref_records, dim3((uint32_t)((n_nodes + REF_NODES - 1) / REF_NODES)),
                     dim3(BLOCK), 0, s, (const float*)nodes.origin, (const float*)nodes.width,
                     n_nodes, mesh_node_base, n_meshes, (uint32_t*)out.verts,
                     (uint32_t*)out.edges);
}

void launch_pack_aos(const NodeArrays& nodes, uint64_t first, uint64_t n, void* dst,
                     hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_pack_aos, dim3(grid_for(42 * n)), dim3(BLOCK), 0, s, nodes, first, n,
                     (uint32_t*)dst);
}

void launch_query_neighbors(const uint32_t* neighbors, uint64_t n_nodes, const uint32_t* node,
                            const uint8_t* slot, uint64_t n, uint32_t* out, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_query, dim3(grid_for(n)), dim3(BLOCK), 0, s, neighbors, n_nodes, node,
                     slot, n, out);
}

#define SVO_INST(K)                                                                              \
  template void launch_emit_level_nodes<K>(int, KeyLayout<K>, const void*, uint64_t,             \
                                           const void*, const uint32_t*, const uint32_t*,        \
                                           uint64_t, const LevelSpan*, uint32_t,                 \
                                           const MeshTreeState*, const MeshNodeLayout*,          \
                                           const MeshParams*, NodeArrays, hipStream_t);          \
  template void launch_emit_roots<K>(KeyLayout<K>, void* const*, uint32_t, const LevelSpan*,     \
                                     uint32_t, const MeshTreeState*, const MeshNodeLayout*,      \
                                     const MeshParams*, int, NodeArrays, hipStream_t);           \
  template void launch_copy_voxels<K>(KeyLayout<K>, const void*, uint64_t, void*, int,           \
                                      hipStream_t);                                              \
  template void launch_neighbors_level<K>(int, KeyLayout<K>, const void*, const uint32_t*,       \
                                          uint64_t, const LevelSpan*, uint32_t,                  \
                                          const MeshTreeState*, const MeshNodeLayout*,           \
                                          NodeArrays, hipStream_t);
SVO_INST(uint32_t)
SVO_INST(uint64_t)
#undef SVO_INST

}  // namespace svo
